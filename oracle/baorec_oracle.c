/*
 * CPU oracle for the BAOrec.jl reconstruction hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load this library, and only as the checker or
 * as the timed CPU restatement; the HIP product path never calls it.
 *
 * It restates, in plain C, the loops of the reference that are not FFTs:
 *   src/mas.jl:1-52      cic!            -> oracle_cic_{f32,f64}
 *   src/mas.jl:218-318   read_cic!       -> oracle_read_cic_{f32,f64}
 *   src/utils.jl:21-25   x_vec           -> oracle_xvec_{f32,f64}
 *   src/multigrid.jl     jacobi!, residual!, prolong!, reduce!, vcycle!, fmg
 *                        -> oracle_mg_*_{f32,f64}
 * The FFT-based parts (smooth!, iterate!, compute_displacements) are restated
 * in numpy in baorec_oracle.py, which drives this library through ctypes.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define REAL float
#define SFX _f32
#include "oracle_impl.h"
#undef REAL
#undef SFX

#define REAL double
#define SFX _f64
#include "oracle_impl.h"
#undef REAL
#undef SFX
