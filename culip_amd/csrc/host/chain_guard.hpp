// Admission of single-launch triangular solves.
//
// A single-launch solve (k_trsv_chain) needs every one of its waves resident
// on the GPU at the same time: its waves wait for values other waves
// publish. Each solve takes a quarter of the wave slots its register use
// leaves on the chip, so at most TRSV_SINGLE_MAX_CONCURRENT of them may run
// on one device together; a fifth could find part of its waves unscheduled
// behind the others and wait until its bounded polls give up.
//
// The library admits a single-launch solve on stream st only while fewer
// than that many *other* streams of the same device still have a
// single-launch solve in flight (launched and not known to be complete).
// Launches on one stream run one after the other and count once. Progress
// is tracked with an event recorded after each admitted launch; a refused
// caller runs the level-scheduled solve instead, which computes the same
// bits. Solves launched by other processes are not seen; streams under
// graph capture are always refused (the captured level launches have no
// residency requirement at replay).
#pragma once

#include <cstdint>
#include <vector>

namespace culip {

// The counting policy, independent of HIP so that it can be unit-tested on
// the host: Stream and Event are opaque handles; `done(ev)` says whether the
// work before ev has completed.
template<typename Stream, typename Event>
class ChainAdmission {
public:
    struct Entry {
        int device;
        Stream stream;
        Event event;
    };

    explicit ChainAdmission(int limit) : limit_(limit) {}

    // Drops completed entries (their events go to `released`) and says
    // whether a solve on (device, stream) may start now.
    template<typename Done>
    bool may_start(int device, Stream stream, Done done, std::vector<Event>& released)
    {
        int others = 0;
        size_t w = 0;
        for(size_t r = 0; r < pending_.size(); ++r) {
            const Entry& e = pending_[r];
            if(done(e.event)) {
                released.push_back(e.event);
                continue;
            }
            if(e.device == device && !(e.stream == stream)) ++others;
            pending_[w++] = e;
        }
        pending_.resize(w);
        return others < limit_;
    }

    // Records that a solve was launched on (device, stream), followed by ev.
    // The stream's earlier entry (completed no later than ev) is replaced and
    // its event goes to `released`.
    void started(int device, Stream stream, Event ev, std::vector<Event>& released)
    {
        for(Entry& e : pending_)
            if(e.device == device && e.stream == stream) {
                released.push_back(e.event);
                e.event = ev;
                return;
            }
        pending_.push_back(Entry{device, stream, ev});
    }

    size_t in_flight() const { return pending_.size(); }
    int limit() const { return limit_; }

private:
    int limit_;
    std::vector<Entry> pending_;
};

}  // namespace culip
