// Device memory comes from the library's stream-ordered pool (cfd_alloc /
// cfd_free); this header keeps the reference's Allocator name
// (include/pool_allocator.hpp:29-33) for code that checks it.
#pragma once

class Allocator {
public:
    static bool allocator_alive;
};
