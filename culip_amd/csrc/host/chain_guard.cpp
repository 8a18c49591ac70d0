// Process-wide admission of single-launch triangular solves (the rule is
// stated in chain_guard.hpp), backed by HIP events.
#include "chain_guard.hpp"

#include <mutex>

#include "objects.hpp"

namespace culip {

namespace {

struct ChainGuard {
    std::mutex mu;
    ChainAdmission<hipStream_t, hipEvent_t> adm{TRSV_SINGLE_MAX_CONCURRENT};
    std::vector<hipEvent_t> spare;   // events free for the next launch
};

// Never destroyed: its events must not be released by a static destructor
// that may run after the HIP runtime has shut down.
ChainGuard& chain_guard()
{
    static ChainGuard* g = new ChainGuard;
    return *g;
}

}  // namespace

bool chain_admitted_launch(hipStream_t st, hipError_t (*launch)(void*), void* ctx)
{
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if(hipStreamIsCapturing(st, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone)
        return false;
    int dev = 0;
    HIP_CHECK(hipGetDevice(&dev));
    ChainGuard& g = chain_guard();
    std::lock_guard<std::mutex> lk(g.mu);
    std::vector<hipEvent_t> released;
    // an event that cannot be queried any more (a failed device) counts as
    // complete, so that a broken stream never blocks admission for good
    const bool ok = g.adm.may_start(dev, st,
        [](hipEvent_t e) { return hipEventQuery(e) != hipErrorNotReady; }, released);
    g.spare.insert(g.spare.end(), released.begin(), released.end());
    if(!ok) return false;
    hipEvent_t ev = nullptr;
    if(g.spare.empty()) {
        HIP_CHECK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    } else {
        ev = g.spare.back();
        g.spare.pop_back();
    }
    const hipError_t e = launch(ctx);
    if(e != hipSuccess) {
        g.spare.push_back(ev);
        hip_check(e, "single-launch triangular solve");
    }
    HIP_CHECK(hipEventRecord(ev, st));
    released.clear();
    g.adm.started(dev, st, ev, released);
    g.spare.insert(g.spare.end(), released.begin(), released.end());
    return true;
}

}  // namespace culip
