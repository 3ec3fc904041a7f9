// honei/util/tags.hh -- backend tags of the MI355X build of HONEI's solve path.
//
// Mirrors the tag scheme of util/tags.hh:31-284: an operator is a class
// template Op<Tag_> whose static value() is specialised per backend.  This
// build adds tags::GPU::HIP (gfx950 kernels behind include/honei_hip.h); the
// CPU tag exists so that memory locks can name host memory.
#pragma once

#include <string>

namespace honei {

enum LockMode { lm_read_only = 0, lm_write_only, lm_read_and_write };

namespace tags {

enum TagValue { tv_cpu = 1, tv_gpu_hip = 64, tv_none = 0 };

struct CPU {
    static const TagValue tag_value = tv_cpu;
    static const TagValue memory_value = tv_cpu;
    static const char *name() { return "cpu"; }
};

struct GPU {
    // gfx950 (MI355X) backend: HIP kernels, one device, stream-ordered.
    struct HIP {
        static const TagValue tag_value = tv_gpu_hip;
        static const TagValue memory_value = tv_gpu_hip;
        static const char *name() { return "hip"; }
    };
};

}  // namespace tags

namespace methods {
struct NONE {};
struct PROLMAT {};
struct JAC {};
struct FIXED {};
struct MIXED {};
namespace CYCLE {
struct V {
    struct STATIC {};
};
struct W {
    struct STATIC {};
};
namespace F {
struct V {
    struct STATIC {};
};
}  // namespace F
}  // namespace CYCLE
}  // namespace methods

// Norm selectors (la/norm.hh)
enum VectorNormType { vnt_max = 0, vnt_l_one, vnt_l_two, vnt_inf };

namespace io_formats {
struct ELL {};
struct EXP {};
struct DV {};
}  // namespace io_formats

}  // namespace honei
