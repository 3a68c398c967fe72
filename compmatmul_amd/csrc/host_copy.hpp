// Host-side copies between pageable and pinned memory on a few threads.
//
// One thread moves about 25 GB/s between pageable and pinned memory on the
// MI355X host, four about 70: enough to keep a pinned DMA (about 56 GB/s)
// busy.  parallel_copy splits one copy into parts done by the caller and by
// persistent worker threads, and returns when every part is done.  One copy
// runs on the workers at a time; a caller that finds them busy (another
// context's thread) copies alone.  A process forked while the pool exists
// starts a new pool in the child, whose copy has no threads.
#pragma once
#include <cstddef>

namespace cm {

// copy len bytes from src to dst (no overlap), complete on return
void parallel_copy(void* dst, const void* src, size_t len);

}  // namespace cm
