// Driver that runs CTLAB's own in-tree functor headers on seeded inputs and
// prints inputs and outputs, so the oracle's restatement of SART weighting and
// FISTA momentum can be checked against the reference code itself.
// Built by oracle/Makefile (target ref) from the headers where they lie in the
// reference tree; output goes to oracle/_ref/ only.  TEST INFRASTRUCTURE ONLY.
#include <cstdio>
#include <cstdint>
#include <omp.h>
#include <thrust/host_vector.h>

#include "SARTWeighting.h"
#include "FastIterativeShrinkageAlgorithm.h"

static uint64_t st = 0x1234567887654321ull;
static float urand()  // xorshift64*, uniform [0,1)
{
    st ^= st >> 12; st ^= st << 25; st ^= st >> 27;
    return (float)((st * 2685821657736338717ull) >> 40) / 16777216.0f;
}

int main()
{
    const int N = 48;
    thrust::host_vector<float> prj(N), real(N), rowSum(N), vol(N), recon(N), colSum(N), cur(N), las(N);
    for (int i = 0; i < N; ++i) {
        prj[i] = 4.0f * urand() - 2.0f;
        real[i] = 4.0f * urand() - 2.0f;
        vol[i] = 4.0f * urand() - 2.0f;
        recon[i] = 4.0f * urand() - 2.0f;
        cur[i] = 4000.0f * urand() - 2000.0f;
        las[i] = 4000.0f * urand() - 2000.0f;
        rowSum[i] = 3.0f * urand();
        colSum[i] = 3.0f * urand();
    }
    // the epsilon guard (1e-7): exactly at, just above, zero and negative
    const float edge[6] = {0.0f, 1.0e-7f, 1.01e-7f, 5.0e-8f, -1.0f, 2.0e-7f};
    for (int k = 0; k < 6; ++k) { rowSum[k] = edge[k]; colSum[k] = edge[k]; }
    thrust::host_vector<float> prj0 = prj, recon0 = recon, cur0 = cur;
    SART_Weighting<float>::Proj(prj, real, rowSum);
    SART_Weighting<float>::Back(vol, recon, colSum);
    const float t1 = 2.5f, t2 = 3.25f;
    FISTA(las, cur, t1, t2);                 // thrust::host_vector overload, default [-2000, 6000]
    thrust::host_vector<float> cur2 = cur0;
    FISTA(las, cur2, t1, t2, -500.0f, 500.0f);
    printf("{\"N\": %d, \"t1\": %.9g, \"t2\": %.9g,\n", N, t1, t2);
    auto dump = [&](const char* nm, const thrust::host_vector<float>& v, bool last) {
        printf("\"%s\": [", nm);
        for (int i = 0; i < N; ++i) printf("%.9g%s", v[i], i + 1 < N ? ", " : "");
        printf("]%s\n", last ? "" : ",");
    };
    dump("prj_in", prj0, false); dump("real", real, false); dump("rowSum", rowSum, false);
    dump("prj_out", prj, false);
    dump("vol", vol, false); dump("recon_in", recon0, false); dump("colSum", colSum, false);
    dump("recon_out", recon, false);
    dump("cur_in", cur0, false); dump("las", las, false); dump("cur_out", cur, false);
    dump("cur_out_clamped", cur2, true);
    printf("}\n");
    return 0;
}
