// engine_f64.hip -- FP64 instantiation (the reference's default double build).
// Compiled with -ffp-contract=off: every operation rounds as written, which
// keeps the results bit-identical to the CPU restatement of the reference.
#include "engine_impl.hpp"
namespace fm {
EngineBase *make_engine_f64(int device, int *err) { return make_engine<double>(device, err); }
}
