// Host unit test of the admission rule for single-launch triangular solves
// (culip_amd/csrc/host/chain_guard.hpp), with integers standing in for
// streams and events. Exits 0 when every check holds.
#include <cstdio>
#include <set>
#include <vector>

#include "../../culip_amd/csrc/host/chain_guard.hpp"

static int failures = 0;
#define CHECK(c)                                                          \
    do {                                                                  \
        if(!(c)) {                                                        \
            std::printf("FAILED line %d: %s\n", __LINE__, #c);            \
            ++failures;                                                   \
        }                                                                 \
    } while(0)

int main()
{
    culip::ChainAdmission<int, int> adm(4);
    std::set<int> done;
    auto is_done = [&](int ev) { return done.count(ev) > 0; };
    std::vector<int> rel;
    int next_ev = 100;
    auto launch = [&](int dev, int stream) {
        if(!adm.may_start(dev, stream, is_done, rel)) return -1;
        const int ev = next_ev++;
        adm.started(dev, stream, ev, rel);
        return ev;
    };

    // four streams of device 0 are admitted, a fifth is refused
    int ev[8];
    for(int s = 0; s < 4; ++s) {
        ev[s] = launch(0, s);
        CHECK(ev[s] >= 0);
    }
    CHECK(adm.in_flight() == 4);
    CHECK(launch(0, 4) < 0);
    // a stream that already has one in flight is admitted again (its
    // launches run one after the other); its entry is replaced
    rel.clear();
    const int again = launch(0, 2);
    CHECK(again >= 0);
    CHECK(adm.in_flight() == 4);
    CHECK(rel.size() == 1 && rel[0] == ev[2]);
    // another device has its own count
    CHECK(launch(1, 4) >= 0);
    CHECK(launch(1, 5) >= 0);
    // once one stream's solve completes, the fifth stream is admitted
    CHECK(launch(0, 4) < 0);
    done.insert(ev[0]);
    rel.clear();
    const int fifth = launch(0, 4);
    CHECK(fifth >= 0);
    CHECK(rel.size() == 1 && rel[0] == ev[0]);     // the completed event is freed
    CHECK(launch(0, 6) < 0);
    // completing the replaced event alone frees nothing: stream 2's entry
    // now waits on its later launch
    done.insert(ev[2]);
    CHECK(launch(0, 6) < 0);
    done.insert(again);
    CHECK(launch(0, 6) >= 0);
    // everything done: the table empties
    for(int e = 100; e < next_ev; ++e) done.insert(e);
    rel.clear();
    CHECK(adm.may_start(0, 9, is_done, rel));
    CHECK(adm.in_flight() == 0);
    // limit 1: a single admitted stream at a time
    culip::ChainAdmission<int, int> one(1);
    std::set<int> d1;
    auto d1f = [&](int e) { return d1.count(e) > 0; };
    CHECK(one.may_start(0, 1, d1f, rel));
    one.started(0, 1, 7, rel);
    CHECK(!one.may_start(0, 2, d1f, rel));
    CHECK(one.may_start(0, 1, d1f, rel));
    d1.insert(7);
    CHECK(one.may_start(0, 2, d1f, rel));
    if(failures == 0) std::printf("chain admission: all checks passed\n");
    return failures ? 1 : 0;
}
