// MFMA path of the expanded-encoding operator (declarations).
#pragma once
#include "common.hpp"
#include "fft.hpp"

namespace mm {

struct HioMfmaCall {
    const double* rowfac;  // [nrows (padded to 16)][rp]
    const double* redfac;  // [nblk*16][rp]
    int rp;                // padded rank: 4, 8, 12 or 16
    int nrows;             // K (forward) or V (adjoint)
    int64_t nred;          // V (forward) or K (adjoint)
    int nblk;              // ceil(nred / 16)
    int mode;              // 0 forward + maps, 1 forward plain, 2 adjoint
    const float2* in;
    const float2* maps;    // coil maps [V x C] or null
    int C;
    int col0, gw;          // column group
    float post;            // output scale (1/sqrt(V))
    const float* W;        // forward output weights or null
    float2* out;
    void* Bs;              // hio_mfma_bs_bytes(nblk, gw)
    float* part;           // nsplit * nrows * 16*NCT floats
    unsigned* amax;        // one word
    int nsplit, bps;       // reduction splits and blocks per split
    int variant;           // bit 0: phase on FP64 MFMA, bit 1: hardware sin/cos, bit 2: software-pipelined form
};

void hio_mfma_apply(hipStream_t st, const HioMfmaCall& c, int prof_iters, double* prof_ms);
size_t hio_mfma_bs_bytes(int nblk, int gw);
void hio_mfma_split(int nrows, int nblk, int slots, int* nsplit, int* bps);

}  // namespace mm
