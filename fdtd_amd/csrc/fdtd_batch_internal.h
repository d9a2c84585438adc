// Batch internals shared by fdtd_batch.cpp and the server (fdtd_serve.cpp).
#pragma once

#include <map>

#include "fdtd_internal.h"

// First-fit allocator over one device allocation (offsets in bytes).
class Arena {
public:
    void init(size_t size)
    {
        size_ = size;
        used_ = 0;
        free_.clear();
        if (size) free_[0] = size;
    }
    bool alloc(size_t len, size_t *off);
    void release(size_t off, size_t len);
    size_t largest() const;
    size_t used() const { return used_; }
    size_t size() const { return size_; }

private:
    size_t size_ = 0, used_ = 0;
    std::map<size_t, size_t> free_;  // offset -> length
};

size_t batch_largest_free(fdtd_batch *b);
size_t batch_used_bytes(fdtd_batch *b);
long long batch_step(fdtd_batch *b);
// Frees a grid's region without waiting; the caller guarantees that no
// enqueued work runs the grid any more (it has reached its step count
// before the last table upload).
int batch_remove_nosync(fdtd_batch *b, int id);
// Enqueues the copy of a grid's first `rows` probe rows to host memory.
int batch_probe_copy_async(fdtd_batch *b, int id, double *host, int rows, hipStream_t stream);
