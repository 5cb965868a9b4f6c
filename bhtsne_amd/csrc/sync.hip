// Cache maintenance where one stream consumes what another stream (or a copy engine) produced.
//
// gfx950 has eight XCDs, each with an L2 of its own that is not coherent with the other seven for
// ordinary device memory: an agent-scope release is a write-back of the L2 of the XCD that
// executes it (buffer_wbl2), an agent-scope acquire an invalidation of that L2 (buffer_inv).
// Within one stream the packet processor brackets every kernel with both.  Where work moves
// between streams (the repulsion forked onto a second stream and joined back, a job prepared on
// its own stream and then advanced by a batch's stream, a batch member finished on its job's
// stream) this library does not rely on the runtime's handling of the cross-stream wait: the
// consuming stream first runs k_l2_sync, whose workgroups go to every XCD (the dispatcher deals
// workgroups to the XCDs in turn), and each performs an agent-scope acquire-release fence there.
// Every XCD's dirty lines are then in memory, and no XCD holds a line older than memory.
#include "internal.hpp"

namespace bht {

namespace {

// 8 workgroups per XCD.
constexpr int L2_SYNC_BLOCKS = 64;

// Release (write-back), wait until it has completed, then acquire (invalidate) and wait, so
// that the kernel ends only after both have taken effect on this XCD.
__global__ void __launch_bounds__(64) k_l2_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    __builtin_amdgcn_s_waitcnt(0);
}

}  // namespace

void l2_sync(hipStream_t s) {
    hipLaunchKernelGGL(k_l2_sync, dim3(L2_SYNC_BLOCKS), dim3(64), 0, s);
    BHT_LAUNCH_CHECK();
}

}  // namespace bht
