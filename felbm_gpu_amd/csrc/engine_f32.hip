// engine_f32.hip -- FP32 distributions and fields with FP64 reductions (the
// reference's -DFELBM_SINGLE build).  Contraction into FMA is allowed here.
#include "engine_impl.hpp"
namespace fm {
EngineBase *make_engine_f32(int device, int *err) { return make_engine<float>(device, err); }
}
