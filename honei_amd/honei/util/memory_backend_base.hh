// honei/util/memory_backend_base.hh -- the memory-backend interface of
// HONEI's MemoryArbiter (util/memory_backend_base.hh:32-58), kept with the
// same methods and argument meanings so a backend written against it plugs
// into the reference's arbiter unchanged: memid identifies a memory block
// (the host array's id), address its host copy, bytes its size.
#pragma once

namespace honei {

class MemoryBackendBase {
  public:
    virtual ~MemoryBackendBase() {}
    // Device copy of block memid: the existing one, else allocated and
    // filled from `address`.
    virtual void *upload(void *memid, void *address, unsigned long bytes) = 0;
    // Device copy of memid back into the host address it was made from.
    virtual void download(void *memid, void *address, unsigned long bytes) = 0;
    // Device storage for memid (contents undefined when new).
    virtual void *alloc(void *memid, void *address, unsigned long bytes) = 0;
    virtual void free(void *memid) = 0;
    virtual void copy(void *src_id, void *src_address, void *dest_id, void *dest_address, unsigned long bytes) = 0;
    // `bytes` counts the source block.
    virtual void convert_float_double(void *src_id, void *src_address, void *dest_id, void *dest_address,
                                      unsigned long bytes) = 0;
    virtual void convert_double_float(void *src_id, void *src_address, void *dest_id, void *dest_address,
                                      unsigned long bytes) = 0;
    virtual void fill(void *memid, void *address, unsigned long bytes, float proto) = 0;
    virtual void fill(void *memid, void *address, unsigned long bytes, double proto) = 0;
    virtual bool knows(void *memid, void *address) = 0;
};

}  // namespace honei
