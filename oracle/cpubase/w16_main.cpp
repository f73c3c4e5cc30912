// Throughput build of the oracle for the CPU baseline (test infrastructure,
// timed only, never used as a checker): the oracle's sources in one
// translation unit, compiled the way the reference's batched CPU library
// is built (-O3 for the host's vector ISA, FMA contraction on, OpenMP SIMD
// loops), with testshade's batched loop shape: 16-lane batches of
// structure-of-arrays data (BatchedShaderGlobals<16>, batched_shader-
// globals.h:77-140), each batch's lane loop an `omp simd` loop, host
// threads over runs of batches (batched_shade_region, src/testshade/
// testshade.cpp:1826-1920). With contraction on, results differ from the
// oracle's in the last bits.
#include "../noise.cpp"
#include "../gabor.cpp"
#include "../color.cpp"
#include "../mtx.cpp"
#include "../closure.cpp"
#include "../shade.cpp"

extern "C" {

// fbm_cell over n points (P: 3 planes; Cout: 3 planes)
void oref_w16_shade_c1(const float* P, float* Cout, int64_t n, int nthreads)
{
    parallel_chunks(n, nthreads, [&](int64_t b, int64_t e) {
        for (int64_t s = b; s < e; s += 16) {
            const int w = (int)std::min<int64_t>(16, e - s);
            float px[16] = {}, py[16] = {}, pz[16] = {}, ox[16], oy[16], oz[16];
            for (int l = 0; l < w; ++l) {
                px[l] = P[s + l];
                py[l] = P[n + s + l];
                pz[l] = P[2 * n + s + l];
            }
#pragma omp simd simdlen(16)
            for (int l = 0; l < 16; ++l) {
                C1Out o = shade_c1(V3(px[l], py[l], pz[l]));
                ox[l]   = o.Cout.x;
                oy[l]   = o.Cout.y;
                oz[l]   = o.Cout.z;
            }
            for (int l = 0; l < w; ++l) {
                Cout[s + l]         = ox[l];
                Cout[n + s + l]     = oy[l];
                Cout[2 * n + s + l] = oz[l];
            }
        }
    });
}

// layered3 over n points (P, N, wo, wi: 3 planes; out: R, G, B, pdf)
void oref_w16_shade_c2(const float* P, const float* N, const float* wo, const float* wi,
                       const float* dPdx, const float* dPdy, float* out, int64_t n, int nthreads)
{
    C2Uniform U;
    U.dPdx = V3(dPdx[0], dPdx[1], dPdx[2]);
    U.dPdy = V3(dPdy[0], dPdy[1], dPdy[2]);
    colorsys_init(U.cs, "Rec709");
    U.gopt.anisotropic = 1;
    U.gopt.direction   = V3(1.0f, 0.0f, 0.0f);
    U.gopt.bandwidth   = 1.0f;
    U.gopt.impulses    = 16.0f;
    U.gopt.do_filter   = 1;
    parallel_chunks(n, nthreads, [&](int64_t b, int64_t e) {
        for (int64_t s = b; s < e; s += 16) {
            const int w = (int)std::min<int64_t>(16, e - s);
            float in[12][16] = {};
            float res[4][16];
            const float* src[4] = { P, N, wo, wi };
            for (int a = 0; a < 4; ++a)
                for (int c = 0; c < 3; ++c)
                    for (int l = 0; l < w; ++l)
                        in[3 * a + c][l] = src[a][c * n + s + l];
#pragma omp simd simdlen(16)
            for (int l = 0; l < 16; ++l) {
                C2Out o = shade_c2(U, V3(in[0][l], in[1][l], in[2][l]), V3(in[3][l], in[4][l], in[5][l]),
                                   V3(in[6][l], in[7][l], in[8][l]), V3(in[9][l], in[10][l], in[11][l]));
                res[0][l] = o.bs.weight.x;
                res[1][l] = o.bs.weight.y;
                res[2][l] = o.bs.weight.z;
                res[3][l] = o.bs.pdf;
            }
            for (int c = 0; c < 4; ++c)
                for (int l = 0; l < w; ++l)
                    out[c * n + s + l] = res[c][l];
        }
    });
}

}  // extern "C"
