// temporary: 1D/3D drivers not yet built
#include "mlau_internal.h"
namespace mlau {
int run_steps_1d(Handle *, int, bool, double, double) { return (int)hipErrorNotSupported; }
int run_steps_3d(Handle *, int, bool, double, double) { return (int)hipErrorNotSupported; }
int run_cfl_1d(Handle *, double, double) { return (int)hipErrorNotSupported; }
int run_cfl_3d(Handle *, double, double) { return (int)hipErrorNotSupported; }
}
