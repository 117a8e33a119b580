// Host-side compression of the operator for the compressed device layout.
#include "ems_compress.h"

#include <algorithm>

namespace ems {

unsigned host_threads()
{
    const unsigned hc = std::thread::hardware_concurrency();
    return std::max(1u, std::min(16u, hc ? hc : 1u));
}

size_t compress_cells(const unsigned N[3], const float* vv, const float* vi, const float* ii, const float* iv,
                      unsigned long long cs, std::vector<uint16_t>& idx_soa, std::vector<float>& tabV,
                      std::vector<float>& tabI, unsigned nthreads)
{
    const size_t Nx = N[0], Ny = N[1], Nz = N[2];
    const size_t M = Nx * Ny * Nz;   // component stride of the N-I-J-K arrays
    // tuple of cell i (i = (x*Ny + y)*Nz + z): the V half then the I half
    auto get = [&](size_t i, float* t) {
        for (int n = 0; n < 3; ++n) {
            t[2 * n] = vv[n * M + i];
            t[2 * n + 1] = vi[n * M + i];
            t[6 + 2 * n] = ii[n * M + i];
            t[6 + 2 * n + 1] = iv[n * M + i];
        }
    };
    std::vector<uint32_t> idx;
    std::vector<float> uniq;
    const size_t U = dedup_tuples<12>(M, get, idx, uniq, 65536, nthreads);
    if (U == 0) return 0;
    tabV.assign(U * 8, 0.0f);
    tabI.assign(U * 8, 0.0f);
    for (size_t u = 0; u < U; ++u)
        for (int k = 0; k < 6; ++k) {
            tabV[u * 8 + k] = uniq[u * 12 + k];
            tabI[u * 8 + k] = uniq[u * 12 + 6 + k];
        }
    idx_soa.assign(cs, 0);
    detail::run_chunks(Nx, nthreads, [&](unsigned, size_t a, size_t b) {
        for (size_t x = a; x < b; ++x)
            for (size_t y = 0; y < Ny; ++y)
                for (size_t z = 0; z < Nz; ++z)
                    idx_soa[x + Nx * (y + Ny * z)] = (uint16_t)idx[(x * Ny + y) * Nz + z];
    });
    return U;
}

void expand_sse_compressed(const unsigned N[3], const unsigned* op_index, unsigned U, const float* const vv4[3],
                           const float* const vi4[3], const float* const ii4[3], const float* const iv4[3],
                           float* vv, float* vi, float* ii, float* iv, unsigned nthreads)
{
    const size_t Nx = N[0], Ny = N[1], Nz = N[2];
    const size_t M = Nx * Ny * Nz;
    const size_t nV = (Nz + 3) / 4;   // Operator_sse::numVectors
    detail::run_chunks(Nx, nthreads, [&](unsigned, size_t a, size_t b) {
        for (size_t x = a; x < b; ++x)
            for (size_t y = 0; y < Ny; ++y)
                for (size_t z = 0; z < Nz; ++z) {
                    const size_t v = z % nV, l = z / nV;
                    const size_t u = op_index[(x * Ny + y) * Nz + v];
                    const size_t c = (x * Ny + y) * Nz + z;
                    for (int n = 0; n < 3; ++n) {
                        // callers validate u < U
                        vv[n * M + c] = vv4[n][u * 4 + l];
                        vi[n * M + c] = vi4[n][u * 4 + l];
                        ii[n * M + c] = ii4[n][u * 4 + l];
                        iv[n * M + c] = iv4[n][u * 4 + l];
                    }
                }
    });
    (void)U;
}

}  // namespace ems
