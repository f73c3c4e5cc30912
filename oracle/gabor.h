// CPU oracle: Gabor noise (sparse convolution, Lagae et al. 2009/2011).
// TEST INFRASTRUCTURE ONLY (see oracle_core.h).
#pragma once
#include "oracle_core.h"

namespace oref {

// NoiseParams subset used by Gabor (src/liboslexec/oslexec_pvt.h:2615-2631)
struct GaborOpts {
    int anisotropic = 0;  // 0 isotropic, 1 anisotropic, 2 hybrid
    int do_filter   = 1;
    V3 direction    = V3(1.0f, 0.0f, 0.0f);
    float bandwidth = 1.0f;
    float impulses  = 16.0f;
};

// GaborUniformParams (src/liboslnoise/sfm_gabornoise.h:100-135)
struct GaborUniform {
    float a, lambda, sqrt_lambda_inv, radius, radius2, radius3, radius_inv;
    float poisson_g;  // expf(-lambda * radius3), the Poisson threshold
};
GaborUniform gabor_uniform(const GaborOpts& opt);

// scalar_gabor (sfm_gabornoise.h:497-515) for a Dual2 point; dim 1..3
// builds P as (x,0,0), (x,y,0) or (x,y,z) (make_Vec3, wide_gabornoise.h).
D2<float> gabor(int dim, const D2<float>* p, const GaborOpts& opt);
// The same with the generator seed (0 for the scalar noise, 0/1/2 for the
// components of gabor3) and optional periods (pgabor; one float per input
// dimension, null for none).
D2<float> gabor_ex(int dim, const D2<float>* p, const GaborOpts& opt, int seed,
                   const float* period);
// scalar_gabor3 / scalar_pgabor3 (sfm_gabornoise.h:517-640): the three
// components are the scalar evaluation with seeds 0, 1 and 2.
D2<V3> gabor3(int dim, const D2<float>* p, const GaborOpts& opt, const float* period);

}  // namespace oref
