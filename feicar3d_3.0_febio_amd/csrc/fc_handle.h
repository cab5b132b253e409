// fc_handle.h -- internal interface between the solver handles
// (fc_poisson.hip) and the request server (fc_serve.hip): a device memory
// pool that handles can allocate from, creation of a handle on a given
// stream and pool, and the description of a prepared handle that the
// batched kernels read.  Not part of the C-ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/fc_poisson.h"
#include "fc_device.h"

namespace fc {

// One hipMalloc'd block carved first-fit into 256-byte aligned pieces.
// Freed pieces coalesce with their neighbours.  alloc and free may be called
// from several threads.
class Arena {
public:
    int init(size_t bytes);
    void release();
    void *alloc(size_t bytes);        // nullptr when no free piece is large enough
    void free(void *p);
    size_t capacity() const { return cap_; }
    size_t in_use() const { return used_; }
    bool owns(const void *p) const { return base_ && (const char *)p >= base_ && (const char *)p < base_ + cap_; }

private:
    std::mutex mu_;
    char *base_ = nullptr;
    size_t cap_ = 0, used_ = 0;
    std::map<size_t, size_t> free_;              // offset -> size
    std::unordered_map<size_t, size_t> live_;    // offset -> size
};

// Host-side shape of a prepared handle: what the server needs to size the
// batched launches of its request.
struct ReqShape {
    int nlev = 0;
    LevelDims d[MAX_LEVELS];
    LineSweep sw[MAX_LEVELS][3];
    std::vector<int> glevel;          // GCM topological level offsets
    int nrb = 0;                      // blocks of the row-strided reductions
    int cycle = 1;                    // V-cycle shape (1 reference, 2 + finest post-smoothing)
    bool mixed = false;
};

// fc_pson_create with the handle's device memory taken from pool (if not
// null) and its work queued on stream (if not null; the handle then does not
// own it).
int handle_create(const fc_grid *gr, const fc_pson_params *p, Arena *pool, hipStream_t stream, fc_pson **out);
// Queue the handle's future work on st (the caller owns st).
void handle_set_stream(fc_pson *h, hipStream_t st);
// Build what one accelerated outer iteration needs (GCM schedule, the
// mixed-precision hierarchy, the Anderson state) and describe the handle:
// desc is the device view (host copy), shape the launch shapes.  Fails when
// the handle's parameters need a host decision inside the iteration.
int handle_describe(fc_pson *h, ReqDesc *desc, ReqShape *shape);
// Device bytes a handle of this shape takes once described (an upper bound
// that admission compares with the pool).
size_t handle_footprint(const fc_grid *gr, const fc_pson_params *p, int nGhost, int nDead);
// the calling thread's last error, as fc_pson_last_error reports it
int set_error(int code, const std::string &msg);

}  // namespace fc
