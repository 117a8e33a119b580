// MI355X (gfx950) EC-FDTD engine: device state, update kernels, extension
// kernels, probes and the C ABI declared in include/ems_hip.h.
//
// Device layout: every field / coefficient array is structure-of-arrays,
// one float32 plane-major block per component, x fastest:
//     a[c * cs + x + Nx * (y + Ny * z)]
// with cs (component stride) rounded up to 64 elements.
//
// Arithmetic follows the reference element update exactly
//     v *= vv;  v += vi * (((a - b) - c) + d)
// (FDTD/engine.cpp:133-168), with separate rounding of every product and sum
// (no contraction: -ffp-contract=off and the pragma below), so results are
// bit-identical to the reference's basic engine on the same operator.
#include <hip/hip_runtime.h>
#include <rocprofiler-sdk-roctx/roctx.h>

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <unordered_map>
#include <thread>
#include <numeric>
#include <vector>

#include "ems_compress.h"
#include "ems_hip.h"

#pragma clang fp contract(off)

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

// a roctx range over a scope (shown by rocprofv3 --marker-trace; free when
// no profiler is attached)
struct TxRange {
    explicit TxRange(const char* m) { roctxRangePushA(m); }
    ~TxRange() { roctxRangePop(); }
    TxRange(const TxRange&) = delete;
    TxRange& operator=(const TxRange&) = delete;
};

#define HIPCHK(expr)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess) return fail(EMS_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

constexpr unsigned kBlock = 256;      // threads per block of the stencil kernels (4 wave64)
constexpr unsigned kXCD = 8;          // accelerator complex dies on MI355X
constexpr unsigned kMaxSlabs = 6;     // UPML slabs (Operator_Ext_UPML::Create_UPML creates at most 6)
constexpr size_t kMaxGraphs = 16;     // distinct timestep-chunk graphs kept per engine

struct GridDesc {
    unsigned Nx, Ny, Nz;
    unsigned plane;     // Nx*Ny
    unsigned long long cs;  // component stride
    unsigned zdepth;    // z-sweep depth per block
    unsigned nzc;       // number of z chunks
    unsigned qtiles;    // (x, y) tiles per z chunk
    unsigned tx;        // tile shape: 0 = kBlock consecutive plane positions, else tx x (kBlock/tx) columns
    unsigned ntx;       // tiles along x (2D tiles)
};

// tiles of one z chunk for a tile shape
inline void set_tiles(GridDesc& g, unsigned tx)
{
    g.tx = tx;
    if (tx == 0) {
        g.ntx = 0;
        g.qtiles = (g.plane + kBlock - 1) / kBlock;
    } else {
        const unsigned ty = kBlock / tx;
        g.ntx = (g.Nx + tx - 1) / tx;
        g.qtiles = g.ntx * ((g.Ny + ty - 1) / ty);
    }
    g.nzc = (g.Nz + g.zdepth - 1) / g.zdepth;
}

// (x, y) column of this thread in tile qt, false outside the mesh
__device__ __forceinline__ bool tile_column(const GridDesc& g, unsigned qt, unsigned& x, unsigned& y)
{
    if (g.tx == 0) {
        const unsigned q = qt * kBlock + threadIdx.x;
        y = q / g.Nx;
        x = q - y * g.Nx;
        return q < g.plane;
    }
    const unsigned ty = kBlock / g.tx;
    const unsigned bx = qt % g.ntx, by = qt / g.ntx;
    x = bx * g.tx + threadIdx.x % g.tx;
    y = by * ty + threadIdx.x / g.tx;
    return x < g.Nx && y < g.Ny;
}

// Coefficients of one half step.  Plain layout: SoA arrays cc (vv or ii) and
// cv (vi or iv).  Compressed layout: one uint16 per cell (shared by both half
// steps) into a table of two float4 per entry, {c_x, v_x, c_y, v_y} and
// {c_z, v_z, 0, 0}, one table per half step.
struct CoefArgs {
    const float* cc;
    const float* cv;
    const unsigned short* idx;
    const float4* tab;
    unsigned ntab;                 // float4 entries of tab (2 per unique tuple)
};

// Kernels that stage a small compressed table in LDS first copy it there
// cooperatively (every thread of the workgroup takes part, then a barrier)
// and then read it instead of the global table.
constexpr unsigned kLdsTabBytes = 32768;    // both half steps' tables of the quad sweep

__device__ __forceinline__ void stage_table(float4* dst, const float4* src, unsigned n)
{
    for (unsigned i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
}

// One UPML slab (Operator_Ext_UPML range).  The positions of all slabs are
// numbered consecutively (slab after slab in creation order, x fastest inside
// a slab); every UPML array is [component or coefficient][position].  Along x
// the numbering is padded to whole groups of four cells: it starts at
// xo = x0 rounded down to a multiple of 4 and spans nx (a multiple of 4)
// columns, and every slab starts at a position that is a multiple of 4, so the
// four cells x = 4k .. 4k+3 of a row have four consecutive positions starting
// at a multiple of 4 (16-byte accesses in the quad kernels).  Padding
// positions belong to no cell.
struct SlabTab {
    int x0, x1, y0, y1, z0, z1;    // inclusive mesh ranges
    int xo;                        // x of the numbering's first column (x0 & ~3)
    unsigned nx, ny;               // numbering extents along x (padded) and y
    unsigned long long off;        // number of the slab's first position (a multiple of 4)
};

// UPML of one half step.  Plain UPML: coef[9][npos] = {c0, cfo, cfn} x
// {x, y, z}.  Compressed UPML: pidx[npos] into a table of three float4 per
// entry, {c0_x, c0_y, c0_z, cfo_x}, {cfo_y, cfo_z, cfn_x, cfn_y},
// {cfn_z, 0, 0, 0}.  Flux [3][npos] is read from flux_in and written to
// flux_out (the same buffer when the half step updates in place).  The slab
// table travels in the kernel arguments.
struct PmlArgs {
    int n;                         // number of slabs (0: no UPML)
    unsigned long long npos;
    SlabTab tab[kMaxSlabs];
    const float* coef;
    const unsigned* pidx;
    const float4* ptab;
    const float* flux_in;
    float* flux_out;
};

enum PmlMode { kNoPml = 0, kPmlPlain = 1, kPmlCompressed = 2 };

// Workgroups are dispatched round-robin over the XCDs by linear block id;
// remap so that each XCD owns one contiguous range of tiles (neighbouring
// rows and z-chunks share that XCD's L2).
__device__ __forceinline__ unsigned xcd_tile(unsigned b, unsigned nb)
{
    const unsigned per = nb / kXCD, rem = nb % kXCD;
    const unsigned xcd = b % kXCD, k = b / kXCD;
    return xcd < rem ? xcd * (per + 1) + k : rem * (per + 1) + (xcd - rem) * per + k;
}

// UPML position of cell (x, y, z), or false outside every slab (slabs do not
// overlap; positions fit 32 bits, checked on the host).  The slab is selected
// first and its position computed once.
__device__ __forceinline__ bool pml_at(const PmlArgs& P, int x, int y, int z, unsigned long long& li)
{
    int sel = -1;
#pragma unroll
    for (int s = 0; s < (int)kMaxSlabs; ++s) {
        const SlabTab& S = P.tab[s];
        if (s < P.n && x >= S.x0 && x <= S.x1 && y >= S.y0 && y <= S.y1 && z >= S.z0 && z <= S.z1) sel = s;
    }
    int bx = 0, by = 0, bz = 0;
    unsigned nx = 0, ny = 0, off = 0;
#pragma unroll
    for (int s = 0; s < (int)kMaxSlabs; ++s) {
        if (sel == s) {
            const SlabTab& S = P.tab[s];
            bx = S.xo; by = S.y0; bz = S.z0;
            nx = S.nx; ny = S.ny; off = (unsigned)S.off;
        }
    }
    li = off + (unsigned)(x - bx) + nx * ((unsigned)(y - by) + ny * (unsigned)(z - bz));
    return sel >= 0;
}

// slabs whose (x, y) range holds column (x, y), one bit per slab
__device__ __forceinline__ unsigned pml_colmask(const PmlArgs& P, int x, int y)
{
    unsigned m = 0;
#pragma unroll
    for (int s = 0; s < (int)kMaxSlabs; ++s) {
        const SlabTab& S = P.tab[s];
        if (s < P.n && x >= S.x0 && x <= S.x1 && y >= S.y0 && y <= S.y1) m |= 1u << s;
    }
    return m;
}

// UPML position of (x, y, z) for a column whose slab mask is m; 0 (a valid
// position) when the cell lies in no slab
__device__ __forceinline__ bool pml_at_col(const PmlArgs& P, unsigned m, int x, int y, int z, unsigned& li)
{
    int sel = -1;
#pragma unroll
    for (int s = 0; s < (int)kMaxSlabs; ++s) {
        const SlabTab& S = P.tab[s];
        if ((m >> s & 1u) && z >= S.z0 && z <= S.z1) sel = s;
    }
    int bx = 0, by = 0, bz = 0;
    unsigned nx = 0, ny = 0, off = 0;
#pragma unroll
    for (int s = 0; s < (int)kMaxSlabs; ++s) {
        if (sel == s) {
            const SlabTab& S = P.tab[s];
            bx = S.xo; by = S.y0; bz = S.z0;
            nx = S.nx; ny = S.ny; off = (unsigned)S.off;
        }
    }
    const unsigned l = off + (unsigned)(x - bx) + nx * ((unsigned)(y - by) + ny * (unsigned)(z - bz));
    li = sel >= 0 ? l : 0u;
    return sel >= 0;
}

// main coefficients of cell p
template <bool CC>
__device__ __forceinline__ void load_coef(const CoefArgs& k, unsigned long long cs, unsigned long long p, unsigned u,
                                          float a[3], float b[3])
{
    if (CC) {
        const float4 t0 = k.tab[2 * u], t1 = k.tab[2 * u + 1];
        a[0] = t0.x; b[0] = t0.y;
        a[1] = t0.z; b[1] = t0.w;
        a[2] = t1.x; b[2] = t1.y;
    } else {
#pragma unroll
        for (int n = 0; n < 3; ++n) {
            a[n] = k.cc[n * cs + p];
            b[n] = k.cv[n * cs + p];
        }
    }
}

// Update of the three components of one cell from its old values x0.
//   plain: f = x0*cc + cv*curl;  out = f
//   UPML : f = F*cc + cv*curl;   out = (c0*x0 - cfo*F) + cfn*f;  F <- f
// (F: the UPML flux).  The UPML lanes add exec-masked loads of the flux and
// of their coefficients or table index.  main = false selects the UPML hooks
// alone, out = (c0*x0 - cfo*F) + cfn*F, for a UPML cell the main update
// skips (the caller then stores no flux).
template <int PM>
__device__ __forceinline__ void cell_math(const float x0[3], const float a[3], const float b[3], const float curl[3],
                                          bool pml, const PmlArgs& P, unsigned long long li, bool store_flux,
                                          float out[3], bool main = true)
{
    float k0[3], ko[3], kn[3], fl[3];
    if (PM != kNoPml && pml) {
#pragma unroll
        for (int n = 0; n < 3; ++n) fl[n] = P.flux_in[n * P.npos + li];
    }
    if (PM == kPmlPlain && pml) {
#pragma unroll
        for (int n = 0; n < 3; ++n) {
            k0[n] = P.coef[n * P.npos + li];
            ko[n] = P.coef[(3 + n) * P.npos + li];
            kn[n] = P.coef[(6 + n) * P.npos + li];
        }
    }
    if (PM == kPmlCompressed && pml) {
        const unsigned u = P.pidx[li];
        const float4 t0 = P.ptab[3 * u], t1 = P.ptab[3 * u + 1], t2 = P.ptab[3 * u + 2];
        k0[0] = t0.x; k0[1] = t0.y; k0[2] = t0.z;
        ko[0] = t0.w; ko[1] = t1.x; ko[2] = t1.y;
        kn[0] = t1.z; kn[1] = t1.w; kn[2] = t2.x;
    }
    float fnew[3];
#pragma unroll
    for (int n = 0; n < 3; ++n) {
        const float xin = (PM != kNoPml && pml) ? fl[n] : x0[n];
        float f = xin * a[n];
        f = f + b[n] * curl[n];
        fnew[n] = f;
        if (PM != kNoPml && pml) {
            // a UPML cell the main update skips takes the hooks alone:
            // out = (c0*x0 - cfo*F) + cfn*F (pml_hooks_only)
            const float t = k0[n] * x0[n] - ko[n] * fl[n];
            out[n] = t + kn[n] * (main ? f : fl[n]);
        } else {
            out[n] = f;
        }
    }
    if (PM != kNoPml && pml && store_flux) {
#pragma unroll
        for (int n = 0; n < 3; ++n) P.flux_out[n * P.npos + li] = fnew[n];
    }
}

// UPML hooks alone, for a UPML position the main update skips (the last
// current line of a direction): Engine_Ext_UPML's pre and post hooks run over
// the whole slab (engine_ext_upml.cpp:135-234), so there
//   out = (c0*x0 - cfo*F) + cfn*F,  F unchanged.
template <int PM>
__device__ __forceinline__ void pml_hooks_only(const float x0[3], const PmlArgs& P, unsigned long long li, float out[3])
{
    float k0[3], ko[3], kn[3], fl[3];
#pragma unroll
    for (int n = 0; n < 3; ++n) fl[n] = P.flux_in[n * P.npos + li];
    if (PM == kPmlPlain) {
#pragma unroll
        for (int n = 0; n < 3; ++n) {
            k0[n] = P.coef[n * P.npos + li];
            ko[n] = P.coef[(3 + n) * P.npos + li];
            kn[n] = P.coef[(6 + n) * P.npos + li];
        }
    } else {
        const unsigned u = P.pidx[li];
        const float4 t0 = P.ptab[3 * u], t1 = P.ptab[3 * u + 1], t2 = P.ptab[3 * u + 2];
        k0[0] = t0.x; k0[1] = t0.y; k0[2] = t0.z;
        ko[0] = t0.w; ko[1] = t1.x; ko[2] = t1.y;
        kn[0] = t1.z; kn[1] = t1.w; kn[2] = t2.x;
    }
#pragma unroll
    for (int n = 0; n < 3; ++n) {
        const float t = k0[n] * x0[n] - ko[n] * fl[n];
        out[n] = t + kn[n] * fl[n];
    }
}

// Field accesses of the cell being updated.  It is read and written once per
// launch; on meshes whose fields exceed the last-level cache a non-temporal
// access keeps the caches for the neighbour rows of the other field, which
// adjacent columns and z chunks read again (256^3 PEC: 0.127 -> 0.103 ms per
// half step).  On smaller meshes the fields stay cached from one half step to
// the next, and plain accesses are faster.
template <bool NT>
__device__ __forceinline__ float ld_own(const float* q)
{
    if (NT) return __builtin_nontemporal_load(q);
    return *q;
}

// a cell's coefficient index, read once per half step: non-temporal with
// the own field (256^3 UPML: 0.3361 -> 0.3350 ms per step)
template <bool NT>
__device__ __forceinline__ unsigned cidx(const unsigned short* q)
{
    if (NT) return (unsigned)__builtin_nontemporal_load(q);
    return (unsigned)*q;
}

template <bool NT>
__device__ __forceinline__ void st_own(float* q, float v)
{
    if (NT) __builtin_nontemporal_store(v, q);
    else *q = v;
}

// Excitation entry as the fused step applies it: cell, component, amplitude
// and delay of one Operator_Ext_Excitation list element.
struct ExcEnt {
    unsigned short x, y, z, dir;
    float amp;
    unsigned delay;
};

// sample of the excitation signal at timestep ts for a delay
// (Engine_Ext_Excitation::Apply2Voltages, engine_ext_excitation.cpp:33-61)
__device__ __forceinline__ int exc_sample(int ts, unsigned delay, int period, unsigned length)
{
    const int p = period > 0 ? period : ts + 1;
    int e = ts - (int)delay;
    e *= (e > 0);
    e %= p;
    e *= (e < (int)length);
    return e;
}

struct ExcArgs {
    const unsigned* beg;           // [ntiles + 1] entry ranges per xy tile, each sorted by z
    const ExcEnt* ent;
    const float* sig;
    unsigned length;
    int period;
    const unsigned* tmask;         // k_volt: bit per tile of kBlock columns holding a source (null: every lane looks)
};

// A quad's source entries ordered by plane (list order within a plane):
// [k, e) are the entries not yet applied, nz the plane of entry k (~0u when
// none is left); a sweep that visits planes in increasing order applies each
// plane's entries once and otherwise only compares nz.
struct SrcCursor {
    unsigned k, e, nz;
};

__device__ __forceinline__ SrcCursor src_cursor(const ExcArgs& ex, unsigned b, unsigned e, unsigned z0)
{
    SrcCursor c{b, e, ~0u};
    while (c.k < e && ex.ent[c.k].z < z0) ++c.k;
    if (c.k < e) c.nz = ex.ent[c.k].z;
    return c;
}

// the entries of plane z added to components `mask` of f, in list order
__device__ __forceinline__ void quad_sources_at(const ExcArgs& ex, SrcCursor& c, unsigned x0, unsigned z,
                                                unsigned mask, int ts, float4 f[3])
{
    if (c.nz != z) return;
    unsigned k = c.k;
    for (; k < c.e; ++k) {
        const ExcEnt E = ex.ent[k];
        if (E.z != z) break;
        if (!(mask >> E.dir & 1u)) continue;
        const float sv = E.amp * ex.sig[exc_sample(ts, E.delay, ex.period, ex.length)];
        const unsigned j = (unsigned)E.x - x0;
#pragma unroll
        for (int n = 0; n < 3; ++n) {
            if (E.dir != n) continue;
            if (j == 0) f[n].x = f[n].x + sv;
            else if (j == 1) f[n].y = f[n].y + sv;
            else if (j == 2) f[n].z = f[n].z + sv;
            else f[n].w = f[n].w + sv;
        }
    }
    c.k = k;
    c.nz = k < c.e ? ex.ent[k].z : ~0u;
}

// the same for one cell's column (f: its three components)
__device__ __forceinline__ void col_sources_at(const ExcArgs& ex, SrcCursor& c, unsigned z, int ts, float f[3])
{
    if (c.nz != z) return;
    unsigned k = c.k;
    for (; k < c.e; ++k) {
        const ExcEnt E = ex.ent[k];
        if (E.z != z) break;
        const float sv = E.amp * ex.sig[exc_sample(ts, E.delay, ex.period, ex.length)];
        if (E.dir == 0) f[0] = f[0] + sv;
        else if (E.dir == 1) f[1] = f[1] + sv;
        else f[2] = f[2] + sv;
    }
    c.k = k;
    c.nz = k < c.e ? ex.ent[k].z : ~0u;
}

// ---------------------------------------------------------------------------
// voltage half step (Engine::UpdateVoltages, FDTD/engine.cpp:110-168) with the
// UPML pre/post hooks of Engine_Ext_UPML (engine_ext_upml.cpp:52-124) fused in:
//   pre : t = vv_u*V - vvfo*F;  V <- F
//   main: V <- V*vv + vi*curl    (computed on the flux)
//   post: F <- V;  V <- t + vvfn*F
// The lower neighbour at index 0 is the cell itself (shift = bool(pos)).
// Each thread owns one (x, y) column and sweeps g.zdepth planes, carrying the
// z-1 currents in registers.  In place.
// ---------------------------------------------------------------------------
template <int PM, bool CC, bool NT>
__device__ __forceinline__ void volt_tile(const GridDesc& g, unsigned t, float* V, const float* __restrict__ I,
                                          const CoefArgs& k, const PmlArgs& pml, const ExcArgs& ex,
                                          const unsigned* __restrict__ numts)
{
    const unsigned qt = t % g.qtiles, zc = t / g.qtiles;
    unsigned x, y;
    if (!tile_column(g, qt, x, y)) return;
    const unsigned z0 = zc * g.zdepth;
    const unsigned z1 = min(z0 + g.zdepth, g.Nz);
    const unsigned long long cs = g.cs;
    const unsigned xm = x ? 1u : 0u, ym = y ? g.Nx : 0u;
    unsigned long long p = x + (unsigned long long)g.Nx * y + (unsigned long long)g.plane * z0;

    const float* __restrict__ Ix = I;
    const float* __restrict__ Iy = I + cs;
    const float* __restrict__ Iz = I + 2 * cs;
    // z-1 values carried in registers along the sweep (self at z = 0)
    const unsigned long long pz = z0 ? p - g.plane : p;
    float Ix_zm = Ix[pz], Iy_zm = Iy[pz];
    // UPML slabs that hold this column (the per-plane test is then along z only)
    const unsigned pcm = PM != kNoPml ? pml_colmask(pml, (int)x, (int)y) : 0u;
    // the column's voltage sources (ordered by plane) where this kernel adds them
    SrcCursor cv{0u, 0u, ~0u};
    int ts = 0;
    if (ex.beg && (!ex.tmask || (ex.tmask[qt >> 5] >> (qt & 31u) & 1u))) {
        cv = src_cursor(ex, ex.beg[x + g.Nx * y], ex.beg[x + g.Nx * y + 1], z0);
        ts = (int)numts[0];
    }

    for (unsigned z = z0; z < z1; ++z, p += g.plane) {
        const float ix = Ix[p], iy = Iy[p], iz = Iz[p];
        const float iz_ym = Iz[p - ym], ix_ym = Ix[p - ym];
        const float iz_xm = Iz[p - xm], iy_xm = Iy[p - xm];
        float curl[3];
        curl[0] = iz - iz_ym - iy + Iy_zm;
        curl[1] = ix - Ix_zm - iz + iz_xm;
        curl[2] = iy - iy_xm - ix + ix_ym;
        Ix_zm = ix;
        Iy_zm = iy;
        float x0[3], a[3], b[3], out[3];
#pragma unroll
        for (int n = 0; n < 3; ++n) x0[n] = ld_own<NT>(&V[n * cs + p]);
        load_coef<CC>(k, cs, p, CC ? cidx<NT>(k.idx + p) : 0u, a, b);
        unsigned li = 0;
        const bool in_pml = PM != kNoPml && pcm && pml_at_col(pml, pcm, (int)x, (int)y, (int)z, li);
        cell_math<PM>(x0, a, b, curl, in_pml, pml, li, true, out);
        col_sources_at(ex, cv, z, ts, out);
#pragma unroll
        for (int n = 0; n < 3; ++n) st_own<NT>(&V[n * cs + p], out[n]);
    }
}

template <int PM, bool CC, bool NT>
__global__ __launch_bounds__(kBlock) void k_volt(const GridDesc g, float* V, const float* __restrict__ I,
                                                 const CoefArgs k, const PmlArgs pml, const ExcArgs ex,
                                                 const unsigned* __restrict__ numts)
{
    volt_tile<PM, CC, NT>(g, xcd_tile(blockIdx.x, gridDim.x), V, I, k, pml, ex, numts);
}

// ---------------------------------------------------------------------------
// current half step (Engine::UpdateCurrents, FDTD/engine.cpp:170-222), same
// UPML fusion with the current flux (engine_ext_upml.cpp:135-234).  Updates
// x < Nx-1, y < Ny-1, z < Nz-1, in place.  Thread (0,0) of tile 0 advances the
// device timestep counter (++numTS of Engine::IterateTS): no kernel of this
// launch reads it, and every later kernel sees the new value.
// ---------------------------------------------------------------------------
template <int PM, bool CC, bool NT>
__device__ __forceinline__ void curr_tile(const GridDesc& g, unsigned t, float* I, const float* __restrict__ V,
                                          const CoefArgs& k, const PmlArgs& pml)
{
    const unsigned qt = t % g.qtiles, zc = t / g.qtiles;
    unsigned x, y;
    if (!tile_column(g, qt, x, y)) return;
    // columns / planes the main update skips only matter for UPML positions
    const bool col_upd = x + 1 < g.Nx && y + 1 < g.Ny;
    if (PM == kNoPml && !col_upd) return;
    const unsigned z0 = zc * g.zdepth;
    const unsigned z1 = min(z0 + g.zdepth, PM == kNoPml ? g.Nz - 1 : g.Nz);
    if (z0 >= z1) return;
    const unsigned long long cs = g.cs;
    // Cells the main update skips (last line of a direction) take the UPML
    // hooks alone, on the same path as the main update: their neighbour
    // loads read the cell itself (valid addresses, values unused), and the
    // result is selected per lane, so a wave holding such a lane does not
    // run a second, serialised branch.
    const unsigned long long dx = col_upd ? 1ull : 0ull, dy = col_upd ? (unsigned long long)g.Nx : 0ull;
    unsigned long long p = x + (unsigned long long)g.Nx * y + (unsigned long long)g.plane * z0;

    const float* __restrict__ Vx = V;
    const float* __restrict__ Vy = V + cs;
    const float* __restrict__ Vz = V + 2 * cs;
    // values at z carried from the previous iteration's z+1 loads
    float vx = Vx[p], vy = Vy[p];
    // UPML slabs that hold this column (the per-plane test is then along z only)
    const unsigned pcm = PM != kNoPml ? pml_colmask(pml, (int)x, (int)y) : 0u;

    for (unsigned z = z0; z < z1; ++z, p += g.plane) {
        const bool main = col_upd && z + 1 < g.Nz;
        const unsigned long long pzp = z + 1 < g.Nz ? p + g.plane : p;
        const float vz = Vz[p];
        const float vx_zp = Vx[pzp], vy_zp = Vy[pzp];
        const float vz_yp = Vz[p + dy], vx_yp = Vx[p + dy];
        const float vz_xp = Vz[p + dx], vy_xp = Vy[p + dx];
        float curl[3];
        curl[0] = vz - vz_yp - vy + vy_zp;
        curl[1] = vx - vx_zp - vz + vz_xp;
        curl[2] = vy - vy_xp - vx + vx_yp;
        vx = vx_zp;
        vy = vy_zp;
        float x0[3], a[3], b[3], out[3];
#pragma unroll
        for (int n = 0; n < 3; ++n) x0[n] = ld_own<NT>(&I[n * cs + p]);
        load_coef<CC>(k, cs, p, CC ? cidx<NT>(k.idx + p) : 0u, a, b);
        unsigned li = 0;
        const bool in_pml = PM != kNoPml && pcm && pml_at_col(pml, pcm, (int)x, (int)y, (int)z, li);
        cell_math<PM>(x0, a, b, curl, in_pml, pml, li, main, out, main);
        if (main || in_pml) {
#pragma unroll
            for (int n = 0; n < 3; ++n) st_own<NT>(&I[n * cs + p], out[n]);
        }
    }
}

template <int PM, bool CC, bool NT>
__global__ __launch_bounds__(kBlock) void k_curr(const GridDesc g, float* I, const float* __restrict__ V,
                                                 const CoefArgs k, const PmlArgs pml, unsigned* __restrict__ numts)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(numts, 1u);
    curr_tile<PM, CC, NT>(g, xcd_tile(blockIdx.x, gridDim.x), I, V, k, pml);
}

// ---------------------------------------------------------------------------
// Quad half steps, for meshes with Nx % 4 == 0: each lane owns four
// consecutive x cells (x0 = 4k .. 4k+3) of one row and sweeps z, so every
// field, coefficient and UPML access is 16 bytes per lane (the x-1 / x+1
// neighbour quads are dword-aligned 16-byte loads at p -+ 1).  Per element the
// arithmetic is volt_tile / curr_tile's, in the same order, so the results are
// bit-identical.  A workgroup covers kBlock quads of one z chunk.
// ---------------------------------------------------------------------------
struct alignas(4) F4u {          // four floats at a dword-aligned address
    float v[4];
};

__device__ __forceinline__ float4 ld4u(const float* q)
{
    F4u t;
    __builtin_memcpy(&t, q, 16);
    return make_float4(t.v[0], t.v[1], t.v[2], t.v[3]);
}

typedef float v4f __attribute__((ext_vector_type(4)));

template <bool NT>
__device__ __forceinline__ float4 ld4(const float* q)
{
    if (NT) {
        const v4f t = __builtin_nontemporal_load(reinterpret_cast<const v4f*>(q));
        return make_float4(t.x, t.y, t.z, t.w);
    }
    return *reinterpret_cast<const float4*>(q);
}

template <bool NT>
__device__ __forceinline__ void st4(float* q, float4 v)
{
    if (NT) {
        const v4f t = {v.x, v.y, v.z, v.w};
        __builtin_nontemporal_store(t, reinterpret_cast<v4f*>(q));
    } else {
        *reinterpret_cast<float4*>(q) = v;
    }
}

__device__ __forceinline__ float f4g(const float4& v, int j) { return j == 0 ? v.x : j == 1 ? v.y : j == 2 ? v.z : v.w; }

__device__ __forceinline__ void f4s(float4& v, int j, float a)
{
    if (j == 0) v.x = a;
    else if (j == 1) v.y = a;
    else if (j == 2) v.z = a;
    else v.w = a;
}

// Boxes of quads a quad half-step launch covers: quads qa .. qa+nq-1 of rows
// ya .. ya+ny-1, planes za .. zb-1, in tiles of kBlock quads (ntiles per z
// chunk of zd planes); `first` is the box's first workgroup.  One box is the whole mesh; a
// UPML mesh splits into the box without UPML cells and the boxes around it.
constexpr int kMaxQBoxes = 8;
struct QBox {
    unsigned qa, nq, ya, ny, za, zb, zd, ntiles, first;   // zd: planes per workgroup
};
// the boxes of one launch in device memory (read with a workgroup-uniform
// index: a by-value array indexed at run time would go to per-lane scratch)
struct QBoxes {
    const QBox* b;
    unsigned n;
};

// quad (x0..x0+3, y) and z range [z0, z1) of lane threadIdx.x in workgroup t;
// false for a lane past its box
__device__ __forceinline__ bool qbox_cell(const QBoxes B, unsigned t, unsigned& x0, unsigned& y, unsigned& z0,
                                          unsigned& z1)
{
    unsigned k = 0;
    for (unsigned i = 1; i < B.n; ++i)
        if (t >= B.b[i].first) k = i;
    const QBox q = B.b[k];
    const unsigned local = t - q.first;
    const unsigned qt = local % q.ntiles, zc = local / q.ntiles;
    const unsigned L = qt * kBlock + threadIdx.x;
    if (L >= q.nq * q.ny) return false;
    const unsigned r = L / q.nq;
    x0 = 4u * (q.qa + (L - r * q.nq));
    y = q.ya + r;
    z0 = q.za + zc * q.zd;
    z1 = min(z0 + q.zd, q.zb);
    return true;
}

// per slab s, bits 4s..4s+3: the elements of quad (x0..x0+3, y) inside the
// slab's x and y ranges
__device__ __forceinline__ unsigned quad_colbits(const PmlArgs& P, int x0, int y)
{
    unsigned m = 0;
#pragma unroll
    for (int s = 0; s < (int)kMaxSlabs; ++s) {
        const SlabTab& S = P.tab[s];
        if (s < P.n && y >= S.y0 && y <= S.y1) {
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if (x0 + j >= S.x0 && x0 + j <= S.x1) m |= 1u << (4 * s + j);
        }
    }
    return m;
}

// UPML part of a quad at plane z.  For every slab holding elements of the
// quad: load the quad's flux (and compressed indices) at the slab's aligned
// position, apply cell_math's UPML formulas to the slab's elements (main: the
// element takes the main update; otherwise the hooks alone), and store the
// flux back (elements outside the slab sit on padding positions and keep
// their values).  in: f[n] = the main update computed on the old values,
// which UPML elements recompute on the flux.
template <int PM, bool STORE_ALL>
__device__ __forceinline__ void quad_pml(const PmlArgs& P, unsigned colbits, int x0, int y, int z, const float4 x0v[3],
                                         const float4 a[3], const float4 b[3], const float4 curl[3], unsigned mainm,
                                         float4 out[3])
{
#pragma unroll
    for (int s = 0; s < (int)kMaxSlabs; ++s) {
        const SlabTab& S = P.tab[s];
        const unsigned bits = (s < P.n && z >= S.z0 && z <= S.z1) ? (colbits >> (4 * s)) & 15u : 0u;
        if (!bits) continue;
        const unsigned li = (unsigned)S.off + (unsigned)(x0 - S.xo) +
                            S.nx * ((unsigned)(y - S.y0) + S.ny * (unsigned)(z - S.z0));
        float4 F[3], Fn[3];
#pragma unroll
        for (int n = 0; n < 3; ++n) {
            F[n] = *reinterpret_cast<const float4*>(P.flux_in + n * P.npos + li);
            Fn[n] = F[n];
        }
        uint4 u4 = make_uint4(0, 0, 0, 0);
        if (PM == kPmlCompressed) u4 = *reinterpret_cast<const uint4*>(P.pidx + li);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (!(bits >> j & 1u)) continue;
            float k0[3], ko[3], kn[3];
            if (PM == kPmlPlain) {
#pragma unroll
                for (int n = 0; n < 3; ++n) {
                    k0[n] = P.coef[n * P.npos + li + j];
                    ko[n] = P.coef[(3 + n) * P.npos + li + j];
                    kn[n] = P.coef[(6 + n) * P.npos + li + j];
                }
            } else {
                const unsigned u = j == 0 ? u4.x : j == 1 ? u4.y : j == 2 ? u4.z : u4.w;
                const float4 t0 = P.ptab[3 * u], t1 = P.ptab[3 * u + 1], t2 = P.ptab[3 * u + 2];
                k0[0] = t0.x; k0[1] = t0.y; k0[2] = t0.z;
                ko[0] = t0.w; ko[1] = t1.x; ko[2] = t1.y;
                kn[0] = t1.z; kn[1] = t1.w; kn[2] = t2.x;
            }
            const bool main = mainm >> j & 1u;
#pragma unroll
            for (int n = 0; n < 3; ++n) {
                const float fl = f4g(F[n], j), xo = f4g(x0v[n], j);
                float f = fl * f4g(a[n], j);
                f = f + f4g(b[n], j) * f4g(curl[n], j);
                const float t = k0[n] * xo - ko[n] * fl;
                f4s(out[n], j, t + kn[n] * (main ? f : fl));
                if (STORE_ALL || main) f4s(Fn[n], j, f);
            }
        }
#pragma unroll
        for (int n = 0; n < 3; ++n) *reinterpret_cast<float4*>(P.flux_out + n * P.npos + li) = Fn[n];
    }
}

// main coefficients of the quad at p (p % 4 == 0)
template <bool CC>
__device__ __forceinline__ void quad_coef(const CoefArgs& k, unsigned long long cs, unsigned long long p, float4 a[3],
                                          float4 b[3])
{
    if (CC) {
        const uint2 w = *reinterpret_cast<const uint2*>(k.idx + p);
        const unsigned u[4] = {w.x & 0xFFFFu, w.x >> 16, w.y & 0xFFFFu, w.y >> 16};
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float4 t0 = k.tab[2 * u[j]], t1 = k.tab[2 * u[j] + 1];
            f4s(a[0], j, t0.x); f4s(b[0], j, t0.y);
            f4s(a[1], j, t0.z); f4s(b[1], j, t0.w);
            f4s(a[2], j, t1.x); f4s(b[2], j, t1.y);
        }
    } else {
#pragma unroll
        for (int n = 0; n < 3; ++n) {
            a[n] = *reinterpret_cast<const float4*>(k.cc + n * cs + p);
            b[n] = *reinterpret_cast<const float4*>(k.cv + n * cs + p);
        }
    }
}

template <int PM, bool CC, bool NT, bool LT>
__global__ __launch_bounds__(kBlock) void k_volt4(const GridDesc g, const QBoxes boxes, float* V,
                                                  const float* __restrict__ I, const CoefArgs kg, const PmlArgs pml,
                                                  const ExcArgs ex, const unsigned* __restrict__ numts)
{
    extern __shared__ float4 s_tab[];
    CoefArgs k = kg;
    if (LT) {
        stage_table(s_tab, kg.tab, kg.ntab);
        __syncthreads();
        k.tab = s_tab;
    }
    unsigned x0, y, z0, z1;
    if (!qbox_cell(boxes, xcd_tile(blockIdx.x, gridDim.x), x0, y, z0, z1)) return;
    const unsigned long long cs = g.cs;
    unsigned long long p = x0 + (unsigned long long)g.Nx * y + (unsigned long long)g.plane * z0;
    const unsigned ym = y ? g.Nx : 0u;
    const unsigned colbits = PM != kNoPml ? quad_colbits(pml, (int)x0, (int)y) : 0u;
    const float* __restrict__ Ix = I;
    const float* __restrict__ Iy = I + cs;
    const float* __restrict__ Iz = I + 2 * cs;
    const unsigned long long pz = z0 ? p - g.plane : p;
    float4 Ix_zm = *reinterpret_cast<const float4*>(Ix + pz), Iy_zm = *reinterpret_cast<const float4*>(Iy + pz);
    // voltage sources added here (per-quad lists ordered by plane) where no
    // hook sits between the update and the excitation; else by k_excite
    SrcCursor cv{0u, 0u, ~0u};
    int ts = 0;
    if (ex.beg) {
        const unsigned qi = x0 / 4 + (g.Nx / 4) * y;
        cv = src_cursor(ex, ex.beg[qi], ex.beg[qi + 1], z0);
        ts = (int)numts[0];
    }
    for (unsigned z = z0; z < z1; ++z, p += g.plane) {
        const float4 ix = *reinterpret_cast<const float4*>(Ix + p);
        const float4 iy = *reinterpret_cast<const float4*>(Iy + p);
        const float4 iz = *reinterpret_cast<const float4*>(Iz + p);
        const float4 iz_ym = *reinterpret_cast<const float4*>(Iz + p - ym);
        const float4 ix_ym = *reinterpret_cast<const float4*>(Ix + p - ym);
        // x-1 neighbours: the quad at p-1, or at x0 = 0 the cell itself for element 0
        float4 iz_xm, iy_xm;
        if (x0) {
            iz_xm = ld4u(Iz + p - 1);
            iy_xm = ld4u(Iy + p - 1);
        } else {
            iz_xm = make_float4(iz.x, iz.x, iz.y, iz.z);
            iy_xm = make_float4(iy.x, iy.x, iy.y, iy.z);
        }
        float4 curl[3];
        curl[0].x = iz.x - iz_ym.x - iy.x + Iy_zm.x;
        curl[0].y = iz.y - iz_ym.y - iy.y + Iy_zm.y;
        curl[0].z = iz.z - iz_ym.z - iy.z + Iy_zm.z;
        curl[0].w = iz.w - iz_ym.w - iy.w + Iy_zm.w;
        curl[1].x = ix.x - Ix_zm.x - iz.x + iz_xm.x;
        curl[1].y = ix.y - Ix_zm.y - iz.y + iz_xm.y;
        curl[1].z = ix.z - Ix_zm.z - iz.z + iz_xm.z;
        curl[1].w = ix.w - Ix_zm.w - iz.w + iz_xm.w;
        curl[2].x = iy.x - iy_xm.x - ix.x + ix_ym.x;
        curl[2].y = iy.y - iy_xm.y - ix.y + ix_ym.y;
        curl[2].z = iy.z - iy_xm.z - ix.z + ix_ym.z;
        curl[2].w = iy.w - iy_xm.w - ix.w + ix_ym.w;
        Ix_zm = ix;
        Iy_zm = iy;
        float4 xv[3], a[3], b[3], out[3];
#pragma unroll
        for (int n = 0; n < 3; ++n) xv[n] = ld4<NT>(V + n * cs + p);
        quad_coef<CC>(k, cs, p, a, b);
#pragma unroll
        for (int n = 0; n < 3; ++n) {
            out[n].x = xv[n].x * a[n].x; out[n].x = out[n].x + b[n].x * curl[n].x;
            out[n].y = xv[n].y * a[n].y; out[n].y = out[n].y + b[n].y * curl[n].y;
            out[n].z = xv[n].z * a[n].z; out[n].z = out[n].z + b[n].z * curl[n].z;
            out[n].w = xv[n].w * a[n].w; out[n].w = out[n].w + b[n].w * curl[n].w;
        }
        if (PM != kNoPml) quad_pml<PM, true>(pml, colbits, (int)x0, (int)y, (int)z, xv, a, b, curl, 15u, out);
        quad_sources_at(ex, cv, x0, z, 7u, ts, out);
#pragma unroll
        for (int n = 0; n < 3; ++n) st4<NT>(V + n * cs + p, out[n]);
    }
}

template <int PM, bool CC, bool NT, bool LT>
__global__ __launch_bounds__(kBlock) void k_curr4(const GridDesc g, const QBoxes boxes, float* I,
                                                  const float* __restrict__ V, const CoefArgs kg, const PmlArgs pml,
                                                  unsigned* __restrict__ numts)
{
    extern __shared__ float4 s_tab[];
    CoefArgs k = kg;
    if (LT) {
        stage_table(s_tab, kg.tab, kg.ntab);
        __syncthreads();
        k.tab = s_tab;
    }
    unsigned x0, y, z0, z1;
    const bool lane = qbox_cell(boxes, xcd_tile(blockIdx.x, gridDim.x), x0, y, z0, z1);
    // one launch of the half step advances the timestep counter
    if (numts && blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(numts, 1u);
    if (!lane) return;
    const bool yupd = y + 1 < g.Ny;
    // elements the main update covers along x (x < Nx-1)
    const unsigned xmain = x0 + 4 < g.Nx ? 15u : 7u;
    if (PM == kNoPml && !yupd) return;
    if (PM == kNoPml) z1 = min(z1, g.Nz - 1);
    if (z0 >= z1) return;
    const unsigned long long cs = g.cs;
    const unsigned colbits = PM != kNoPml ? quad_colbits(pml, (int)x0, (int)y) : 0u;
    // neighbours of cells the main update skips read the cell itself
    const unsigned long long dy = yupd ? (unsigned long long)g.Nx : 0ull;
    const bool xlast = x0 + 4 >= g.Nx;
    unsigned long long p = x0 + (unsigned long long)g.Nx * y + (unsigned long long)g.plane * z0;
    const float* __restrict__ Vx = V;
    const float* __restrict__ Vy = V + cs;
    const float* __restrict__ Vz = V + 2 * cs;
    float4 vx = *reinterpret_cast<const float4*>(Vx + p), vy = *reinterpret_cast<const float4*>(Vy + p);
    for (unsigned z = z0; z < z1; ++z, p += g.plane) {
        const bool zupd = z + 1 < g.Nz;
        const unsigned mainm = (yupd && zupd) ? xmain : 0u;
        const unsigned long long pzp = zupd ? p + g.plane : p;
        const float4 vz = *reinterpret_cast<const float4*>(Vz + p);
        const float4 vx_zp = *reinterpret_cast<const float4*>(Vx + pzp);
        const float4 vy_zp = *reinterpret_cast<const float4*>(Vy + pzp);
        const float4 vz_yp = *reinterpret_cast<const float4*>(Vz + p + dy);
        const float4 vx_yp = *reinterpret_cast<const float4*>(Vx + p + dy);
        // x+1 neighbours: the quad at p+1, or in the last quad of a row the
        // cell itself for element 3 (its value is not used)
        float4 vz_xp, vy_xp;
        if (!xlast) {
            vz_xp = ld4u(Vz + p + 1);
            vy_xp = ld4u(Vy + p + 1);
        } else {
            vz_xp = make_float4(vz.y, vz.z, vz.w, vz.w);
            vy_xp = make_float4(vy.y, vy.z, vy.w, vy.w);
        }
        float4 curl[3];
        curl[0].x = vz.x - vz_yp.x - vy.x + vy_zp.x;
        curl[0].y = vz.y - vz_yp.y - vy.y + vy_zp.y;
        curl[0].z = vz.z - vz_yp.z - vy.z + vy_zp.z;
        curl[0].w = vz.w - vz_yp.w - vy.w + vy_zp.w;
        curl[1].x = vx.x - vx_zp.x - vz.x + vz_xp.x;
        curl[1].y = vx.y - vx_zp.y - vz.y + vz_xp.y;
        curl[1].z = vx.z - vx_zp.z - vz.z + vz_xp.z;
        curl[1].w = vx.w - vx_zp.w - vz.w + vz_xp.w;
        curl[2].x = vy.x - vy_xp.x - vx.x + vx_yp.x;
        curl[2].y = vy.y - vy_xp.y - vx.y + vx_yp.y;
        curl[2].z = vy.z - vy_xp.z - vx.z + vx_yp.z;
        curl[2].w = vy.w - vy_xp.w - vx.w + vx_yp.w;
        vx = vx_zp;
        vy = vy_zp;
        float4 xv[3], a[3], b[3], out[3];
#pragma unroll
        for (int n = 0; n < 3; ++n) xv[n] = ld4<NT>(I + n * cs + p);
        quad_coef<CC>(k, cs, p, a, b);
#pragma unroll
        for (int n = 0; n < 3; ++n) {
            float4 f;
            f.x = xv[n].x * a[n].x; f.x = f.x + b[n].x * curl[n].x;
            f.y = xv[n].y * a[n].y; f.y = f.y + b[n].y * curl[n].y;
            f.z = xv[n].z * a[n].z; f.z = f.z + b[n].z * curl[n].z;
            f.w = xv[n].w * a[n].w; f.w = f.w + b[n].w * curl[n].w;
            // cells the main update skips keep their values (UPML ones are set below)
            out[n].x = (mainm & 1u) ? f.x : xv[n].x;
            out[n].y = (mainm & 2u) ? f.y : xv[n].y;
            out[n].z = (mainm & 4u) ? f.z : xv[n].z;
            out[n].w = (mainm & 8u) ? f.w : xv[n].w;
        }
        if (PM != kNoPml) quad_pml<PM, false>(pml, colbits, (int)x0, (int)y, (int)z, xv, a, b, curl, mainm, out);
        if (PM != kNoPml || mainm) {
#pragma unroll
            for (int n = 0; n < 3; ++n) st4<NT>(I + n * cs + p, out[n]);
        }
    }
}

// ---------------------------------------------------------------------------
// One full timestep in one z-sweep (voltage half step, voltage excitation,
// current half step, current excitation, with UPML fused in both halves):
// the same per-element arithmetic as k_volt / k_curr + the excitation hooks,
// so results are bit-identical to the two-half-step path.
//
// Fields and UPML fluxes are double-buffered: the step reads state n (V0, I0,
// flux_in) and writes state n+1 (V1, I1, flux_out), so workgroups never see
// each other's new values.  A workgroup owns a tile of kTX x kTY (x, y)
// columns and a z chunk; each wave owns 4 rows of 64 columns.  Per plane z it
// computes V^{n+1}(z) for its tile plus a halo (column x0+kTX, row y0+kTY) into
// LDS (the halo values are recomputed, not stored), then I^{n+1}(z-1) from the
// LDS planes z-1 and z.  The first plane of a chunk reads I^n(z0-1) from
// memory; the plane z1 after the chunk is recomputed for I^{n+1}(z1-1).
// Current-side sources must lie inside the updated range (x < Nx-1, ...).
// The last workgroup to finish advances the device timestep counter.
// ---------------------------------------------------------------------------
constexpr unsigned kTX = 64, kTY = 16, kRowsPerWave = kTY / (kBlock / 64);

struct StepArgs {
    GridDesc g;
    unsigned ntx, nty;             // xy tiles
    const float* V0;
    const float* I0;
    float* V1;
    float* I1;
    CoefArgs kV, kI;
    PmlArgs pV, pI;
    ExcArgs eV, eI;
    const unsigned* numts;         // [0] timestep counter (advanced by k_tick after the step)
};

template <int PM, bool CC>
__device__ __forceinline__ void volt_cell(const StepArgs& a, unsigned x, unsigned y, unsigned z, unsigned long long p,
                                          unsigned u, float ix, float iy, float iz, float ix_zm, float iy_zm,
                                          bool store_flux, int ts, unsigned eb, unsigned ee, float v[3])
{
    const GridDesc& g = a.g;
    const unsigned long long cs = g.cs;
    const unsigned xm = x ? 1u : 0u, ym = y ? g.Nx : 0u;
    const float* __restrict__ I0 = a.I0;
    const float iz_ym = I0[2 * cs + p - ym], ix_ym = I0[p - ym];
    const float iz_xm = I0[2 * cs + p - xm], iy_xm = I0[cs + p - xm];
    float curl[3];
    curl[0] = iz - iz_ym - iy + iy_zm;
    curl[1] = ix - ix_zm - iz + iz_xm;
    curl[2] = iy - iy_xm - ix + ix_ym;
    float x0[3], ca[3], cb[3];
#pragma unroll
    for (int n = 0; n < 3; ++n) x0[n] = a.V0[n * cs + p];
    load_coef<CC>(a.kV, cs, p, u, ca, cb);
    unsigned long long li = 0;
    const bool in_pml = PM != kNoPml && pml_at(a.pV, (int)x, (int)y, (int)z, li);
    cell_math<PM>(x0, ca, cb, curl, in_pml, a.pV, li, store_flux, v);
    // voltage sources of this cell in list order
    for (unsigned e = eb; e < ee; ++e) {
        const ExcEnt E = a.eV.ent[e];
        if (E.z == z && E.x == x && E.y == y) {
            const int s = exc_sample(ts, E.delay, a.eV.period, a.eV.length);
            v[E.dir] = v[E.dir] + E.amp * a.eV.sig[s];
        }
    }
}

// UPML hooks alone at a current cell the main update skips (fused step):
// I^{n+1} from I^n, the flux unchanged (both flux buffers hold it)
template <int PM>
__device__ __forceinline__ void pml_hooks_cell(const StepArgs& a, unsigned x, unsigned y, unsigned z,
                                               unsigned long long p, const float i0[3])
{
    unsigned long long li = 0;
    if (!pml_at(a.pI, (int)x, (int)y, (int)z, li)) return;
    float out[3];
    pml_hooks_only<PM>(i0, a.pI, li, out);
#pragma unroll
    for (int n = 0; n < 3; ++n) a.I1[n * a.g.cs + p] = out[n];
}

template <int PM, bool CC>
__global__ __launch_bounds__(kBlock) void k_step(const StepArgs a)
{
    __shared__ float sV[2][3][kTY + 1][kTX + 1];
    const GridDesc& g = a.g;
    const unsigned long long cs = g.cs;
    const unsigned t = xcd_tile(blockIdx.x, gridDim.x);
    const unsigned tx = t % a.ntx, ty = (t / a.ntx) % a.nty, zc = t / (a.ntx * a.nty);
    const unsigned tile = tx + a.ntx * ty;
    const unsigned x0 = tx * kTX, y0 = ty * kTY;
    const unsigned z0 = zc * g.zdepth;
    const unsigned z1 = min(z0 + g.zdepth, g.Nz);
    const unsigned zlast = min(z1, g.Nz - 1);
    const unsigned lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const unsigned x = x0 + lane;
    const int ts = (int)a.numts[0];
    const unsigned evb = a.eV.beg[tile], eve = a.eV.beg[tile + 1];
    const unsigned eib = a.eI.beg[tile], eie = a.eI.beg[tile + 1];
    // halo cells: column x0+kTX (rows 0..kTY) and row y0+kTY (columns 0..kTX-1)
    const bool hcol = x0 + kTX <= g.Nx - 1, hrow = y0 + kTY <= g.Ny - 1;
    const unsigned hi = threadIdx.x;
    const bool h_on_col = hi <= kTY;
    const unsigned hx = h_on_col ? x0 + kTX : x0 + (hi - (kTY + 1));
    const unsigned hy = h_on_col ? y0 + hi : y0 + kTY;
    const bool has_halo = h_on_col ? (hcol && hy < g.Ny) : (hi < kTY + 1 + kTX && hrow && hx < g.Nx);
    const unsigned hr = hy - y0, hc = hx - x0;

    // per row: I^n carried along z (x, y components at z-1) and the own I^n
    // and coefficient index at the previous plane for the current update
    float izm_x[kRowsPerWave], izm_y[kRowsPerWave];
    float iown[kRowsPerWave][3], inew[kRowsPerWave][3];
    unsigned uown[kRowsPerWave], unew[kRowsPerWave];
    const unsigned long long pz0 = z0 ? (unsigned long long)(z0 - 1) * g.plane : 0ull;
#pragma unroll
    for (int r = 0; r < (int)kRowsPerWave; ++r) {
        const unsigned y = y0 + wave * kRowsPerWave + r;
        izm_x[r] = izm_y[r] = 0.0f;
        uown[r] = unew[r] = 0;
        iown[r][0] = iown[r][1] = iown[r][2] = 0.0f;
        inew[r][0] = inew[r][1] = inew[r][2] = 0.0f;
        if (x < g.Nx && y < g.Ny) {
            const unsigned long long p = x + (unsigned long long)g.Nx * y + (z0 ? pz0 : (unsigned long long)z0 * g.plane);
            izm_x[r] = a.I0[p];
            izm_y[r] = a.I0[cs + p];
        }
    }
    float h_izm_x = 0.0f, h_izm_y = 0.0f;
    if (has_halo) {
        const unsigned long long p = hx + (unsigned long long)g.Nx * hy + (z0 ? pz0 : 0ull);
        h_izm_x = a.I0[p];
        h_izm_y = a.I0[cs + p];
    }

    for (unsigned z = z0; z <= zlast; ++z) {
        const unsigned s = z & 1u;
        const unsigned long long zp = (unsigned long long)z * g.plane;
        // -- V^{n+1} at plane z: own rows
#pragma unroll
        for (int r = 0; r < (int)kRowsPerWave; ++r) {
            const unsigned row = wave * kRowsPerWave + r;
            const unsigned y = y0 + row;
            if (x < g.Nx && y < g.Ny) {
                const unsigned long long p = x + (unsigned long long)g.Nx * y + zp;
                const float ix = a.I0[p], iy = a.I0[cs + p], iz = a.I0[2 * cs + p];
                const unsigned u = CC ? a.kV.idx[p] : 0u;
                float v[3];
                volt_cell<PM, CC>(a, x, y, z, p, u, ix, iy, iz, izm_x[r], izm_y[r], z < z1, ts, evb, eve, v);
                izm_x[r] = ix;
                izm_y[r] = iy;
                sV[s][0][row][lane] = v[0];
                sV[s][1][row][lane] = v[1];
                sV[s][2][row][lane] = v[2];
                if (z < z1) {
#pragma unroll
                    for (int n = 0; n < 3; ++n) a.V1[n * cs + p] = v[n];
                }
                // keep I^n and the coefficient index of this cell for I^{n+1}(z)
                inew[r][0] = ix; inew[r][1] = iy; inew[r][2] = iz;
                unew[r] = u;
            }
        }
        // -- halo V^{n+1} at plane z (LDS only)
        if (has_halo) {
            const unsigned long long p = hx + (unsigned long long)g.Nx * hy + zp;
            const float ix = a.I0[p], iy = a.I0[cs + p], iz = a.I0[2 * cs + p];
            const unsigned u = CC ? a.kV.idx[p] : 0u;
            float v[3];
            volt_cell<PM, CC>(a, hx, hy, z, p, u, ix, iy, iz, h_izm_x, h_izm_y, false, ts, evb, eve, v);
            h_izm_x = ix;
            h_izm_y = iy;
            sV[s][0][hr][hc] = v[0];
            sV[s][1][hr][hc] = v[1];
            sV[s][2][hr][hc] = v[2];
        }
        __syncthreads();
        // -- I^{n+1} at plane z-1 from the LDS planes z-1 (slot s^1) and z (slot s)
        if (z > z0) {
            const unsigned zc1 = z - 1;
            const unsigned long long zq = (unsigned long long)zc1 * g.plane;
#pragma unroll
            for (int r = 0; r < (int)kRowsPerWave; ++r) {
                const unsigned row = wave * kRowsPerWave + r;
                const unsigned y = y0 + row;
                if (PM != kNoPml && x < g.Nx && y < g.Ny && !(x + 1 < g.Nx && y + 1 < g.Ny)) {
                    pml_hooks_cell<PM>(a, x, y, zc1, x + (unsigned long long)g.Nx * y + zq, iown[r]);
                } else if (x + 1 < g.Nx && y + 1 < g.Ny) {
                    const unsigned long long p = x + (unsigned long long)g.Nx * y + zq;
                    const unsigned o = s ^ 1u;
                    const float vx = sV[o][0][row][lane], vy = sV[o][1][row][lane], vz = sV[o][2][row][lane];
                    const float vx_zp = sV[s][0][row][lane], vy_zp = sV[s][1][row][lane];
                    const float vz_yp = sV[o][2][row + 1][lane], vx_yp = sV[o][0][row + 1][lane];
                    const float vz_xp = sV[o][2][row][lane + 1], vy_xp = sV[o][1][row][lane + 1];
                    float curl[3];
                    curl[0] = vz - vz_yp - vy + vy_zp;
                    curl[1] = vx - vx_zp - vz + vz_xp;
                    curl[2] = vy - vy_xp - vx + vx_yp;
                    float ca[3], cb[3], out[3];
                    load_coef<CC>(a.kI, cs, p, uown[r], ca, cb);
                    unsigned long long li = 0;
                    const bool in_pml = PM != kNoPml && pml_at(a.pI, (int)x, (int)y, (int)zc1, li);
                    cell_math<PM>(iown[r], ca, cb, curl, in_pml, a.pI, li, true, out);
                    for (unsigned e = eib; e < eie; ++e) {
                        const ExcEnt E = a.eI.ent[e];
                        if (E.z == zc1 && E.x == x && E.y == y) {
                            const int sm = exc_sample(ts, E.delay, a.eI.period, a.eI.length);
                            out[E.dir] = out[E.dir] + E.amp * a.eI.sig[sm];
                        }
                    }
#pragma unroll
                    for (int n = 0; n < 3; ++n) a.I1[n * cs + p] = out[n];
                }
            }
        }
        // I^n and the index of plane z serve the current update at the next plane
#pragma unroll
        for (int r = 0; r < (int)kRowsPerWave; ++r) {
            iown[r][0] = inew[r][0]; iown[r][1] = inew[r][1]; iown[r][2] = inew[r][2];
            uown[r] = unew[r];
        }
        __syncthreads();
    }
    // the last current plane is skipped by the main update: UPML hooks only
    if (PM != kNoPml && z1 == g.Nz) {
        const unsigned long long zq = (unsigned long long)(g.Nz - 1) * g.plane;
#pragma unroll
        for (int r = 0; r < (int)kRowsPerWave; ++r) {
            const unsigned y = y0 + wave * kRowsPerWave + r;
            if (x < g.Nx && y < g.Ny) pml_hooks_cell<PM>(a, x, y, g.Nz - 1, x + (unsigned long long)g.Nx * y + zq, iown[r]);
        }
    }
}

// ---------------------------------------------------------------------------
// Fused single-sweep step without LDS or barriers (k_step2).  The same step as
// k_step (state n in V0, I0 and the flux inputs; state n+1 into V1, I1 and the
// flux outputs), on the two-half-step tiles: a workgroup owns kBlock
// consecutive (x, y) columns and a z chunk, and each thread sweeps its own
// column.  The current update at (x, y, z-1) needs V^{n+1} of its column at
// planes z-1 and z and of the columns (x+1, y) (components z, y) and
// (x, y+1) (components z, x) at plane z-1.  The thread recomputes those four
// neighbour components itself, with the same arithmetic as their owners
// (sources included, flux not stored), and carries plane z-1 in registers,
// so no value crosses threads.  Each plane first issues every load (clamped
// to valid addresses where a value is not needed, so the body has no
// branches before its arithmetic), then computes, then stores.  Sources come
// from per-column lists: the entries of column q are beg[q] .. beg[q+1], in
// list order.
// ---------------------------------------------------------------------------
struct Step2Args {
    GridDesc g;
    CoefArgs kV, kI;
    PmlArgs pV, pI;                // one slab table and index array for both halves
    ExcArgs eV, eI;                // beg: [plane + 1] per-column ranges
    const unsigned* numts;         // [0] timestep counter (advanced by k_tick after the step)
};

// main coefficients (c, v) of component n at cell p: plain arrays or table
// entry u, whose float2 n is {c_n, v_n}
template <bool CC>
__device__ __forceinline__ float2 coef_n(const CoefArgs& k, unsigned long long cs, unsigned long long p, unsigned u, int n)
{
    if (CC) return reinterpret_cast<const float2*>(k.tab)[4 * u + n];
    return make_float2(k.cc[n * cs + p], k.cv[n * cs + p]);
}

// UPML coefficients (c0, cfo, cfn) of component n at position li: plain
// arrays or the 12 floats of table entry w
template <int PM>
__device__ __forceinline__ void pml_coef_n(const PmlArgs& P, unsigned li, unsigned w, int n, float& k0, float& ko,
                                           float& kn)
{
    if (PM == kPmlPlain) {
        k0 = P.coef[n * P.npos + li];
        ko = P.coef[(3 + n) * P.npos + li];
        kn = P.coef[(6 + n) * P.npos + li];
    } else {
        const float* t = reinterpret_cast<const float*>(P.ptab) + 12 * (unsigned long long)w;
        k0 = t[n];
        ko = t[3 + n];
        kn = t[6 + n];
    }
}

// one component of the update (cell_math for a single component): returns
// the new value, f receives the new flux
template <int PM>
__device__ __forceinline__ float comp_math(float x0, float2 c, float curl, bool pml, float fl, float k0, float ko, float kn,
                                           float& f)
{
    const float xin = (PM != kNoPml && pml) ? fl : x0;
    float g = xin * c.x;
    g = g + c.y * curl;
    f = g;
    if (PM == kNoPml) return g;
    const float t = k0 * x0 - ko * fl;
    const float o = t + kn * g;
    return pml ? o : g;
}

// sources of one column at plane z and component n, added in list order
// (Engine_Ext_Excitation::Apply2Voltages / Apply2Current)
__device__ __forceinline__ float col_sources(const ExcArgs& e, unsigned b, unsigned end, unsigned z, unsigned n, int ts,
                                             float v)
{
    for (unsigned k = b; k < end; ++k) {
        const ExcEnt E = e.ent[k];
        if (E.z == z && E.dir == n) v = v + E.amp * e.sig[exc_sample(ts, E.delay, e.period, e.length)];
    }
    return v;
}

// sources of cell (x, column group) at plane z and component n, from a list
// holding several columns' entries (in list order per cell)
__device__ __forceinline__ float cell_sources(const ExcArgs& e, unsigned b, unsigned end, unsigned x, unsigned z,
                                              unsigned n, int ts, float v)
{
    for (unsigned k = b; k < end; ++k) {
        const ExcEnt E = e.ent[k];
        if (E.x == x && E.z == z && E.dir == n) v = v + E.amp * e.sig[exc_sample(ts, E.delay, e.period, e.length)];
    }
    return v;
}

template <int PM, bool CC>
__global__ __launch_bounds__(kBlock) void k_step2(const Step2Args a, const float* __restrict__ V0,
                                                  const float* __restrict__ I0, float* __restrict__ V1,
                                                  float* __restrict__ I1, const float* __restrict__ fVin,
                                                  float* __restrict__ fVout, const float* __restrict__ fIin,
                                                  float* __restrict__ fIout)
{
    const GridDesc& g = a.g;
    const unsigned long long cs = g.cs;
    const unsigned t = xcd_tile(blockIdx.x, gridDim.x);
    const unsigned qt = t % g.qtiles, zc = t / g.qtiles;
    unsigned x, y;
    const bool valid = tile_column(g, qt, x, y);
    const int ts = (int)a.numts[0];
    if (valid) {
        const unsigned Nx = g.Nx, Ny = g.Ny, Nz = g.Nz;
        const unsigned long long plane = g.plane;
        const unsigned long long npV = a.pV.npos, npI = a.pI.npos;
        const unsigned z0 = zc * g.zdepth;
        const unsigned z1 = min(z0 + g.zdepth, Nz);
        const unsigned zlast = min(z1, Nz - 1);
        const unsigned q = x + Nx * y;
        const bool upd = x + 1 < Nx && y + 1 < Ny;        // the current update runs in this column
        const unsigned long long xm = x ? 1ull : 0ull, ym = y ? (unsigned long long)Nx : 0ull;
        // neighbour columns (the column itself where no current update runs)
        const unsigned long long dr = upd ? 1ull : 0ull, du = upd ? (unsigned long long)Nx : 0ull;
        // source ranges of the own column and of the two neighbour columns
        const unsigned evb = a.eV.beg[q], eve = a.eV.beg[q + 1];
        const unsigned eib = a.eI.beg[q], eie = a.eI.beg[q + 1];
        const unsigned erb = a.eV.beg[q + dr], ere = upd ? a.eV.beg[q + 2] : erb;
        const unsigned eub = a.eV.beg[q + du], eue = upd ? a.eV.beg[q + Nx + 1] : eub;
        unsigned m_own = 0, m_r = 0, m_u = 0;
        if (PM != kNoPml) {
            m_own = pml_colmask(a.pV, (int)x, (int)y);
            if (upd) {
                m_r = pml_colmask(a.pV, (int)x + 1, (int)y);
                m_u = pml_colmask(a.pV, (int)x, (int)y + 1);
            }
        }
        // I^n at z-1 (the cell itself at z = 0): own x, y; (x+1, y) x; (x, y+1) y
        const unsigned long long pzm = z0 ? (unsigned long long)(z0 - 1) * plane + q : (unsigned long long)q;
        float ix_zm = I0[pzm], iy_zm = I0[cs + pzm];
        float ixr_zm = I0[pzm + dr], iyu_zm = I0[cs + pzm + du];
        // plane z-1 carried for the current update: own V^{n+1} and I^n, the
        // neighbour components, the coefficient index, UPML position and index
        float vm0 = 0.0f, vm1 = 0.0f, vm2 = 0.0f, vrz = 0.0f, vry = 0.0f, vuz = 0.0f, vux = 0.0f;
        float im0 = 0.0f, im1 = 0.0f, im2 = 0.0f;
        unsigned um = 0, lim = 0, wm = 0;
        bool pmlm = false;
        unsigned long long pm = q;
        for (unsigned z = z0; z <= zlast; ++z) {
            const unsigned long long p = (unsigned long long)z * plane + q;
            const unsigned long long pr = p + dr, pu = p + du;
            // ---- loads
            const float ix = I0[p], iy = I0[cs + p], iz = I0[2 * cs + p];
            const float iz_ym = I0[2 * cs + p - ym], ix_ym = I0[p - ym];
            const float iz_xm = I0[2 * cs + p - xm], iy_xm = I0[cs + p - xm];
            const float ixr = I0[pr], iyr = I0[cs + pr], izr = I0[2 * cs + pr], ixr_ym = I0[pr - ym];
            const float ixu = I0[pu], iyu = I0[cs + pu], izu = I0[2 * cs + pu], iyu_xm = I0[cs + pu - xm];
            const float v0x = V0[p], v0y = V0[cs + p], v0z = V0[2 * cs + p];
            const float r0z = V0[2 * cs + pr], r0y = V0[cs + pr];
            const float u0z = V0[2 * cs + pu], u0x = V0[pu];
            const unsigned u = CC ? a.kV.idx[p] : 0u;
            const unsigned ur = CC ? a.kV.idx[pr] : 0u;
            const unsigned uu = CC ? a.kV.idx[pu] : 0u;
            const float2 cV0 = coef_n<CC>(a.kV, cs, p, u, 0), cV1 = coef_n<CC>(a.kV, cs, p, u, 1),
                         cV2 = coef_n<CC>(a.kV, cs, p, u, 2);
            const float2 cR2 = coef_n<CC>(a.kV, cs, pr, ur, 2), cR1 = coef_n<CC>(a.kV, cs, pr, ur, 1);
            const float2 cU2 = coef_n<CC>(a.kV, cs, pu, uu, 2), cU0 = coef_n<CC>(a.kV, cs, pu, uu, 0);
            const float2 cI0 = coef_n<CC>(a.kI, cs, pm, um, 0), cI1 = coef_n<CC>(a.kI, cs, pm, um, 1),
                         cI2 = coef_n<CC>(a.kI, cs, pm, um, 2);
            unsigned li = 0, lr = 0, lu = 0, w = 0;
            bool pml = false, pmr = false, pmu = false;
            float fl[3] = {0.0f, 0.0f, 0.0f}, k0[3] = {0.0f, 0.0f, 0.0f}, ko[3] = {0.0f, 0.0f, 0.0f},
                  kn[3] = {0.0f, 0.0f, 0.0f};
            float flr2 = 0.0f, flr1 = 0.0f, flu2 = 0.0f, flu0 = 0.0f;
            float kr[2][3] = {{0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}}, ku[2][3] = {{0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}};
            float flI[3] = {0.0f, 0.0f, 0.0f}, kI0[3] = {0.0f, 0.0f, 0.0f}, kIo[3] = {0.0f, 0.0f, 0.0f},
                  kIn[3] = {0.0f, 0.0f, 0.0f};
            if (PM != kNoPml) {
                pml = pml_at_col(a.pV, m_own, (int)x, (int)y, (int)z, li);
                pmr = pml_at_col(a.pV, m_r, (int)x + 1, (int)y, (int)z, lr);
                pmu = pml_at_col(a.pV, m_u, (int)x, (int)y + 1, (int)z, lu);
                w = PM == kPmlCompressed ? a.pV.pidx[li] : 0u;
                const unsigned wr = PM == kPmlCompressed ? a.pV.pidx[lr] : 0u;
                const unsigned wu = PM == kPmlCompressed ? a.pV.pidx[lu] : 0u;
#pragma unroll
                for (int n = 0; n < 3; ++n) {
                    fl[n] = fVin[n * npV + li];
                    pml_coef_n<PM>(a.pV, li, w, n, k0[n], ko[n], kn[n]);
                    flI[n] = fIin[n * npI + lim];
                    pml_coef_n<PM>(a.pI, lim, wm, n, kI0[n], kIo[n], kIn[n]);
                }
                flr2 = fVin[2 * npV + lr];
                flr1 = fVin[npV + lr];
                pml_coef_n<PM>(a.pV, lr, wr, 2, kr[0][0], kr[0][1], kr[0][2]);
                pml_coef_n<PM>(a.pV, lr, wr, 1, kr[1][0], kr[1][1], kr[1][2]);
                flu2 = fVin[2 * npV + lu];
                flu0 = fVin[lu];
                pml_coef_n<PM>(a.pV, lu, wu, 2, ku[0][0], ku[0][1], ku[0][2]);
                pml_coef_n<PM>(a.pV, lu, wu, 0, ku[1][0], ku[1][1], ku[1][2]);
            }
            // ---- own V^{n+1}(x, y, z)
            float v[3], fv[3];
            {
                const float c0 = iz - iz_ym - iy + iy_zm;
                const float c1 = ix - ix_zm - iz + iz_xm;
                const float c2 = iy - iy_xm - ix + ix_ym;
                v[0] = comp_math<PM>(v0x, cV0, c0, pml, fl[0], k0[0], ko[0], kn[0], fv[0]);
                v[1] = comp_math<PM>(v0y, cV1, c1, pml, fl[1], k0[1], ko[1], kn[1], fv[1]);
                v[2] = comp_math<PM>(v0z, cV2, c2, pml, fl[2], k0[2], ko[2], kn[2], fv[2]);
            }
            // ---- neighbour components: (x+1, y) z and y, (x, y+1) z and x
            float nrz, nry, nuz, nux, dummy;
            {
                const float c2r = iyr - iy - ixr + ixr_ym;
                const float c1r = ixr - ixr_zm - izr + iz;
                nrz = comp_math<PM>(r0z, cR2, c2r, pmr, flr2, kr[0][0], kr[0][1], kr[0][2], dummy);
                nry = comp_math<PM>(r0y, cR1, c1r, pmr, flr1, kr[1][0], kr[1][1], kr[1][2], dummy);
                const float c2u = iyu - iyu_xm - ixu + ix;
                const float c0u = izu - iz - iyu + iyu_zm;
                nuz = comp_math<PM>(u0z, cU2, c2u, pmu, flu2, ku[0][0], ku[0][1], ku[0][2], dummy);
                nux = comp_math<PM>(u0x, cU0, c0u, pmu, flu0, ku[1][0], ku[1][1], ku[1][2], dummy);
            }
            // ---- voltage sources
            if (evb != eve) {
#pragma unroll
                for (int n = 0; n < 3; ++n) v[n] = col_sources(a.eV, evb, eve, z, n, ts, v[n]);
            }
            if (erb != ere) {
                nrz = col_sources(a.eV, erb, ere, z, 2, ts, nrz);
                nry = col_sources(a.eV, erb, ere, z, 1, ts, nry);
            }
            if (eub != eue) {
                nuz = col_sources(a.eV, eub, eue, z, 2, ts, nuz);
                nux = col_sources(a.eV, eub, eue, z, 0, ts, nux);
            }
            // ---- I^{n+1}(x, y, z-1): main update, or the UPML hooks alone in
            // a column the main update skips
            float io[3], fi[3];
            {
                const float c0 = vm2 - vuz - vm1 + v[1];
                const float c1 = vm0 - v[0] - vm2 + vrz;
                const float c2 = vm1 - vry - vm0 + vux;
                const float im[3] = {im0, im1, im2};
                io[0] = comp_math<PM>(im[0], cI0, c0, pmlm, flI[0], kI0[0], kIo[0], kIn[0], fi[0]);
                io[1] = comp_math<PM>(im[1], cI1, c1, pmlm, flI[1], kI0[1], kIo[1], kIn[1], fi[1]);
                io[2] = comp_math<PM>(im[2], cI2, c2, pmlm, flI[2], kI0[2], kIo[2], kIn[2], fi[2]);
                if (PM != kNoPml && !upd) {
#pragma unroll
                    for (int n = 0; n < 3; ++n) {
                        const float tt = kI0[n] * im[n] - kIo[n] * flI[n];
                        io[n] = tt + kIn[n] * flI[n];
                    }
                }
            }
            const bool wI = z > z0 && (upd || (PM != kNoPml && pmlm));
            if (z > z0 && upd && eib != eie) {
#pragma unroll
                for (int n = 0; n < 3; ++n) io[n] = col_sources(a.eI, eib, eie, z - 1, n, ts, io[n]);
            }
            // ---- stores
            if (z < z1) {
                V1[p] = v[0];
                V1[cs + p] = v[1];
                V1[2 * cs + p] = v[2];
                if (PM != kNoPml && pml) {
                    fVout[li] = fv[0];
                    fVout[npV + li] = fv[1];
                    fVout[2 * npV + li] = fv[2];
                }
            }
            if (wI) {
                I1[pm] = io[0];
                I1[cs + pm] = io[1];
                I1[2 * cs + pm] = io[2];
                if (PM != kNoPml && pmlm && upd) {
                    fIout[lim] = fi[0];
                    fIout[npI + lim] = fi[1];
                    fIout[2 * npI + lim] = fi[2];
                }
            }
            // ---- carry plane z
            vm0 = v[0]; vm1 = v[1]; vm2 = v[2];
            vrz = nrz; vry = nry; vuz = nuz; vux = nux;
            im0 = ix; im1 = iy; im2 = iz;
            ix_zm = ix; iy_zm = iy;
            ixr_zm = ixr; iyu_zm = iyu;
            um = u; lim = li; wm = w; pmlm = pml;
            pm = p;
        }
        // the last current plane is skipped by the main update: UPML hooks only
        if (PM != kNoPml && z1 == Nz && pmlm) {
            const float im[3] = {im0, im1, im2};
#pragma unroll
            for (int n = 0; n < 3; ++n) {
                float c0, co, cn;
                pml_coef_n<PM>(a.pI, lim, wm, n, c0, co, cn);
                const float fl = fIin[n * npI + lim];
                const float tt = c0 * im[n] - co * fl;
                I1[n * cs + pm] = tt + cn * fl;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Quad column sweep: one full timestep per launch, four x cells per lane
// (16-byte accesses), over a box of cells without UPML (k_step2's scheme with
// quads).  Each lane owns the quad x0 .. x0+3 of row y and sweeps a z chunk;
// per plane z it loads state n once and computes
//   * V^{n+1}(x0..x0+3, y, z), stored;
//   * V^{n+1} z and x components of the quad of row y+1 and z and y
//     components of the cell (x0+4, y): the neighbours its current update
//     reads, recomputed (not stored) with their owners' arithmetic, sources
//     included;
//   * I^{n+1}(x0..x0+3, y, z-1) from those, stored.
// The box is [4*qa, 4*(qa+nq)) x [ya, ya+ny) x [za, zb); voltages are stored
// on all of it, currents where x + 1 < xb, y + 1 < yb and z + 1 < zb (the
// rest of the box's current update, and every cell outside the box, belongs
// to another pass or to no update at all).  Per element the arithmetic is
// k_step2's (and so k_volt / k_curr's): results are bit-identical.
// ---------------------------------------------------------------------------
// A Mur face on x = 0 (neighbour line 1) or x = Nx-1 (neighbour Nx-2) whose
// pre and post hooks k_step4t runs (its points are numbered z * Ny + y)
struct MurFold {
    const float* cP;               // coefficients of the y and z components
    const float* cPP;
    float* mP;                     // the hooks' values, read by k_mur_apply
    float* mPP;
    unsigned start_ts;
    unsigned on;
};

struct Step4Args {
    GridDesc g;
    CoefArgs kV, kI;
    ExcArgs eV, eI;                // beg: [plane + 1] per-column ranges in list order
    const unsigned* numts;
    unsigned qa, nq, ya, ny, za, zb, xb, yb;
    unsigned zdepth, ntiles;       // planes per chunk; tiles of kBlock quads over nq * ny
    unsigned ntq;                  // k_step4t: tiles of 64 quads along x (ntiles = ntq x row tiles)
    MurFold mf[2];                 // k_step4t: the x = 0 and x = Nx-1 Mur faces
};

// coefficient pair (c, v) of component n of the quad at p (table entries u[])
template <bool CC>
__device__ __forceinline__ void quad_coef_n(const CoefArgs& k, unsigned long long cs, unsigned long long p,
                                            const unsigned u[4], int n, float4& c, float4& v)
{
    if (CC) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float2 t = reinterpret_cast<const float2*>(k.tab)[4 * u[j] + n];
            f4s(c, j, t.x);
            f4s(v, j, t.y);
        }
    } else {
        c = *reinterpret_cast<const float4*>(k.cc + n * cs + p);
        v = *reinterpret_cast<const float4*>(k.cv + n * cs + p);
    }
}

__device__ __forceinline__ void quad_idx(const unsigned short* idx, unsigned long long p, unsigned u[4])
{
    const uint2 w = *reinterpret_cast<const uint2*>(idx + p);
    u[0] = w.x & 0xFFFFu;
    u[1] = w.x >> 16;
    u[2] = w.y & 0xFFFFu;
    u[3] = w.y >> 16;
}

__device__ __forceinline__ float4 quad_math(float4 x0, float4 c, float4 v, float4 curl)
{
    float4 o;
    o.x = x0.x * c.x; o.x = o.x + v.x * curl.x;
    o.y = x0.y * c.y; o.y = o.y + v.y * curl.y;
    o.z = x0.z * c.z; o.z = o.z + v.z * curl.z;
    o.w = x0.w * c.w; o.w = o.w + v.w * curl.w;
    return o;
}

// a - b - c + d per element, in that order
__device__ __forceinline__ float4 quad_curl(float4 a, float4 b, float4 c, float4 d)
{
    float4 o;
    o.x = a.x - b.x - c.x + d.x;
    o.y = a.y - b.y - c.y + d.y;
    o.z = a.z - b.z - c.z + d.z;
    o.w = a.w - b.w - c.w + d.w;
    return o;
}

// the sources of columns x0 .. x0+3 (entries [b, e), list order) at plane z
// added to components `mask` (bit n: component n) of f
__device__ __forceinline__ void quad_sources(const ExcArgs& ex, unsigned b, unsigned e, unsigned x0, unsigned z,
                                             unsigned mask, int ts, float4 f[3])
{
    for (unsigned k = b; k < e; ++k) {
        const ExcEnt E = ex.ent[k];
        if (E.z != z || !(mask >> E.dir & 1u)) continue;
        const float sv = E.amp * ex.sig[exc_sample(ts, E.delay, ex.period, ex.length)];
        const unsigned j = (unsigned)E.x - x0;
#pragma unroll
        for (int n = 0; n < 3; ++n) {
            if (E.dir != n) continue;
            if (j == 0) f[n].x = f[n].x + sv;
            else if (j == 1) f[n].y = f[n].y + sv;
            else if (j == 2) f[n].z = f[n].z + sv;
            else f[n].w = f[n].w + sv;
        }
    }
}

template <bool CC, bool NT, bool LT>
__global__ __launch_bounds__(kBlock) void k_step4(const Step4Args ag, const float* __restrict__ V0,
                                                  const float* __restrict__ I0, float* __restrict__ V1,
                                                  float* __restrict__ I1)
{
    extern __shared__ float4 s_tab[];
    Step4Args a = ag;
    if (LT) {
        stage_table(s_tab, ag.kV.tab, ag.kV.ntab);
        stage_table(s_tab + ag.kV.ntab, ag.kI.tab, ag.kI.ntab);
        __syncthreads();
        a.kV.tab = s_tab;
        a.kI.tab = s_tab + ag.kV.ntab;
    }
    const GridDesc& g = a.g;
    const unsigned long long cs = g.cs;
    const unsigned t = xcd_tile(blockIdx.x, gridDim.x);
    const unsigned qt = t % a.ntiles, zc = t / a.ntiles;
    const unsigned L = qt * kBlock + threadIdx.x;
    if (L >= a.nq * a.ny) return;
    const unsigned r = L / a.nq;
    const unsigned x0 = 4u * (a.qa + (L - r * a.nq)), y = a.ya + r;
    const int ts = (int)a.numts[0];
    const unsigned Nx = g.Nx;
    const unsigned long long plane = g.plane;
    const unsigned z0 = a.za + zc * a.zdepth;
    const unsigned z1 = min(z0 + a.zdepth, a.zb);
    // the plane after the chunk is recomputed for the currents of its last plane
    const unsigned zlast = z1 < a.zb ? z1 : z1 - 1;
    const unsigned q = x0 + Nx * y;
    const bool yupd = y + 1 < a.yb;                    // the row's current update runs here
    const bool xr = x0 + 4 < a.xb;                     // element 3's current update runs here
    const unsigned imask = !yupd ? 0u : xr ? 15u : (x0 + 3 < a.xb ? 7u : x0 + 2 < a.xb ? 3u : x0 + 1 < a.xb ? 1u : 0u);
    const unsigned long long ym = y ? (unsigned long long)Nx : 0ull;
    // neighbours read by the current update (the cell itself where none runs)
    const unsigned long long du = yupd ? (unsigned long long)Nx : 0ull, dr = xr ? 4ull : 0ull;
    const unsigned evb = a.eV.beg[q], eve = a.eV.beg[q + 4];
    const unsigned eib = a.eI.beg[q], eie = a.eI.beg[q + 4];
    const unsigned eub = a.eV.beg[q + du], eue = yupd ? a.eV.beg[q + du + 4] : eub;
    // cell x0+4: the next quad's list (ordered by plane), entries of x0+4 only
    const unsigned erb = a.eV.beg[q + dr], ere = xr ? a.eV.beg[q + dr + 4] : erb;
    // V^{n+1} of plane z-1 carried for the current update: the own quad, the
    // z and x components of row y+1, the z and y components of (x0+4, y)
    float4 vm[3] = {}, vu2m = {}, vu0m = {};
    float vr2m = 0.0f, vr1m = 0.0f;
    for (unsigned z = z0; z <= zlast; ++z) {
        const unsigned long long p = (unsigned long long)z * plane + q;
        // I^n at z-1 (the cell itself at z = 0); read again where needed
        const unsigned long long pz = z ? p - plane : p;
        // ---- own V^{n+1}(x0..x0+3, y, z)
        const float4 ix = *reinterpret_cast<const float4*>(I0 + p);
        const float4 iy = *reinterpret_cast<const float4*>(I0 + cs + p);
        const float4 iz = *reinterpret_cast<const float4*>(I0 + 2 * cs + p);
        float4 v[3];
        unsigned u[4] = {0u, 0u, 0u, 0u};
        {
            const float4 iz_ym = *reinterpret_cast<const float4*>(I0 + 2 * cs + p - ym);
            const float4 ix_ym = *reinterpret_cast<const float4*>(I0 + p - ym);
            const float4 ix_m = *reinterpret_cast<const float4*>(I0 + pz);
            const float4 iy_m = *reinterpret_cast<const float4*>(I0 + cs + pz);
            float4 iz_xm, iy_xm;
            if (x0) {
                iz_xm = ld4u(I0 + 2 * cs + p - 1);
                iy_xm = ld4u(I0 + cs + p - 1);
            } else {
                iz_xm = make_float4(iz.x, iz.x, iz.y, iz.z);
                iy_xm = make_float4(iy.x, iy.x, iy.y, iy.z);
            }
            float4 v0[3];
            v0[0] = ld4<NT>(V0 + p);
            v0[1] = ld4<NT>(V0 + cs + p);
            v0[2] = ld4<NT>(V0 + 2 * cs + p);
            if (CC) quad_idx(a.kV.idx, p, u);
            float4 c, w;
            quad_coef_n<CC>(a.kV, cs, p, u, 0, c, w);
            v[0] = quad_math(v0[0], c, w, quad_curl(iz, iz_ym, iy, iy_m));
            quad_coef_n<CC>(a.kV, cs, p, u, 1, c, w);
            v[1] = quad_math(v0[1], c, w, quad_curl(ix, ix_m, iz, iz_xm));
            quad_coef_n<CC>(a.kV, cs, p, u, 2, c, w);
            v[2] = quad_math(v0[2], c, w, quad_curl(iy, iy_xm, ix, ix_ym));
            if (evb != eve) quad_sources(a.eV, evb, eve, x0, z, 7u, ts, v);
        }
        if (z < z1) {
#pragma unroll
            for (int n = 0; n < 3; ++n) st4<NT>(V1 + n * cs + p, v[n]);
        }
        // ---- I^{n+1}(x0..x0+3, y, z-1) from the carried plane and v
        if (z > z0 && z < a.zb && imask) {
            const unsigned long long pm = pz;
            unsigned um[4] = {0u, 0u, 0u, 0u};
            if (CC) quad_idx(a.kI.idx, pm, um);
            float4 io[3];
            const float4 vzr = make_float4(vm[2].y, vm[2].z, vm[2].w, vr2m);
            const float4 vyr = make_float4(vm[1].y, vm[1].z, vm[1].w, vr1m);
            float4 c, w;
            quad_coef_n<CC>(a.kI, cs, pm, um, 0, c, w);
            io[0] = quad_math(*reinterpret_cast<const float4*>(I0 + pm), c, w, quad_curl(vm[2], vu2m, vm[1], v[1]));
            quad_coef_n<CC>(a.kI, cs, pm, um, 1, c, w);
            io[1] = quad_math(*reinterpret_cast<const float4*>(I0 + cs + pm), c, w, quad_curl(vm[0], v[0], vm[2], vzr));
            quad_coef_n<CC>(a.kI, cs, pm, um, 2, c, w);
            io[2] = quad_math(*reinterpret_cast<const float4*>(I0 + 2 * cs + pm), c, w,
                              quad_curl(vm[1], vyr, vm[0], vu0m));
            if (eib != eie) quad_sources(a.eI, eib, eie, x0, z - 1, 7u, ts, io);
            if (imask == 15u) {
#pragma unroll
                for (int n = 0; n < 3; ++n) st4<NT>(I1 + n * cs + pm, io[n]);
            } else {
#pragma unroll
                for (int n = 0; n < 3; ++n)
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        if (imask >> j & 1u) I1[n * cs + pm + j] = f4g(io[n], j);
            }
        }
        vm[0] = v[0];
        vm[1] = v[1];
        vm[2] = v[2];
        // ---- neighbours at z: row y+1 (z and x components), cell (x0+4, y) (z
        // and y), for the current update of plane z in the next iteration
        if (z < zlast) {
            const unsigned long long pu = p + du;
            const float4 ixu = *reinterpret_cast<const float4*>(I0 + pu);
            const float4 iyu = *reinterpret_cast<const float4*>(I0 + cs + pu);
            const float4 izu = *reinterpret_cast<const float4*>(I0 + 2 * cs + pu);
            const float4 iyu_m = *reinterpret_cast<const float4*>(I0 + cs + pz + du);
            float4 iyu_xm;
            if (x0) iyu_xm = ld4u(I0 + cs + pu - 1);
            else iyu_xm = make_float4(iyu.x, iyu.x, iyu.y, iyu.z);
            const float4 u0z = *reinterpret_cast<const float4*>(V0 + 2 * cs + pu);
            const float4 u0x = *reinterpret_cast<const float4*>(V0 + pu);
            unsigned uu[4] = {0u, 0u, 0u, 0u};
            if (CC) quad_idx(a.kV.idx, pu, uu);
            float4 nu[3];
            float4 c, w;
            quad_coef_n<CC>(a.kV, cs, pu, uu, 2, c, w);
            nu[2] = quad_math(u0z, c, w, quad_curl(iyu, iyu_xm, ixu, ix));
            quad_coef_n<CC>(a.kV, cs, pu, uu, 0, c, w);
            nu[0] = quad_math(u0x, c, w, quad_curl(izu, iz, iyu, iyu_m));
            nu[1] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (eub != eue) quad_sources(a.eV, eub, eue, x0, z, 5u, ts, nu);
            vu2m = nu[2];
            vu0m = nu[0];
            const unsigned long long pr = p + dr;
            const float ixr = I0[pr], iyr = I0[cs + pr], izr = I0[2 * cs + pr], ixr_ym = I0[pr - ym];
            const float ixr_m = I0[pz + dr];
            const float r0z = V0[2 * cs + pr], r0y = V0[cs + pr];
            const unsigned ur = CC ? a.kV.idx[pr] : 0u;
            const float2 c2 = coef_n<CC>(a.kV, cs, pr, ur, 2), c1 = coef_n<CC>(a.kV, cs, pr, ur, 1);
            const float c2r = iyr - iy.w - ixr + ixr_ym;
            const float c1r = ixr - ixr_m - izr + iz.w;
            float nrz = r0z * c2.x;
            nrz = nrz + c2.y * c2r;
            float nry = r0y * c1.x;
            nry = nry + c1.y * c1r;
            if (erb != ere) {
                nrz = cell_sources(a.eV, erb, ere, x0 + 4, z, 2, ts, nrz);
                nry = cell_sources(a.eV, erb, ere, x0 + 4, z, 1, ts, nry);
            }
            vr2m = nrz;
            vr1m = nry;
        }
    }
}

// ---------------------------------------------------------------------------
// Tiled quad column sweep (form 3 on meshes of at least 64 quads per row).
// A workgroup is a tile of 64 quads x 4 rows, one wave per row, sweeping a
// z chunk in lock step (one barrier per plane).  The V^{n+1} values the
// current update needs from its neighbours come from the cells' owners: the
// z and x components of row y+1 through LDS (three plane buffers, so one
// barrier per plane separates every write from the reads of the same
// buffer), element 0 of quad x0+4 from the next lane of the wave.  Only the
// tile's top row (wave 3) recomputes row y+1 as k_step4 does, and only lane
// 63 recomputes cell (x0+4, y); the x-1 neighbours of the voltage update
// come from the previous lane and I^n of plane z-1 is carried.  Sources are
// per-quad lists ordered by plane, consumed through cursors.  Arithmetic
// is k_step4's; the Mur faces on x = 0 and x = Nx-1 run their pre and post
// hooks here (below), the other faces' hooks around the launch.  Lanes past
// the mesh run the loop on clamped positions and store nothing.
// ---------------------------------------------------------------------------
constexpr unsigned kQTileRows = kBlock / 64;

// Mur pre and post hooks of point k (coefficients cp, cpp) for face element
// F and neighbour element S of a quad (k_mur_pre / k_mur_post arithmetic):
// y and z from V^n (v0) and V^{n+1} (v); stored when `keep`
template <int F, int S>
__device__ __forceinline__ void mur_fold_quad(const MurFold& M, unsigned k, float cp, float cpp, const float4 v0[3],
                                              const float4 v[3], bool keep, float& fy, float& fz)
{
    float my = f4g(v0[1], S) - cp * f4g(v0[1], F);
    my = my + cp * f4g(v[1], S);
    float mz = f4g(v0[2], S) - cpp * f4g(v0[2], F);
    mz = mz + cpp * f4g(v[2], S);
    if (keep) {
        M.mP[k] = my;
        M.mPP[k] = mz;
    }
    fy = my;
    fz = mz;
}

template <bool CC, bool NT, bool LT>
__global__ __launch_bounds__(kBlock, 3) void k_step4t(const Step4Args ag, const float* __restrict__ V0,
                                                   const float* __restrict__ I0, float* __restrict__ V1,
                                                   float* __restrict__ I1)
{
    extern __shared__ float4 s_tab[];
    __shared__ float4 s_vz[3][kBlock], s_vx[3][kBlock];
    Step4Args a = ag;
    if (LT) {
        stage_table(s_tab, ag.kV.tab, ag.kV.ntab);
        stage_table(s_tab + ag.kV.ntab, ag.kI.tab, ag.kI.ntab);
        __syncthreads();
        a.kV.tab = s_tab;
        a.kI.tab = s_tab + ag.kV.ntab;
    }
    const GridDesc& g = a.g;
    const unsigned long long cs = g.cs;
    const unsigned t = xcd_tile(blockIdx.x, gridDim.x);
    const unsigned tt = t % a.ntiles, zc = t / a.ntiles;
    const unsigned lane = threadIdx.x & 63u, row = threadIdx.x >> 6;
    const unsigned tqx = tt % a.ntq, tqy = tt / a.ntq;
    const unsigned qx = tqx * 64u + lane, yy = tqy * kQTileRows + row;
    const bool active = qx < a.nq && yy < a.ny;
    // lanes past the mesh compute on a clamped quad and store nothing
    const unsigned x0 = 4u * min(qx, a.nq - 1u), y = min(yy, a.ny - 1u);
    const int ts = (int)a.numts[0];
    const unsigned Nx = g.Nx;
    const unsigned long long plane = g.plane;
    const unsigned z0 = a.za + zc * a.zdepth;
    const unsigned z1 = min(z0 + a.zdepth, a.zb);
    const unsigned zlast = z1 < a.zb ? z1 : z1 - 1;
    const unsigned q = x0 + Nx * y;
    const bool yupd = y + 1 < a.yb;
    const bool xr = x0 + 4 < a.xb;
    const unsigned imask = !active || !yupd ? 0u
                           : xr ? 15u : (x0 + 3 < a.xb ? 7u : x0 + 2 < a.xb ? 3u : x0 + 1 < a.xb ? 1u : 0u);
    const unsigned long long ym = y ? (unsigned long long)Nx : 0ull;
    const unsigned long long du = yupd ? (unsigned long long)Nx : 0ull, dr = xr ? 4ull : 0ull;
    // the top row of the tile recomputes row y+1; lane 63 recomputes (x0+4, y)
    const bool top = row == kQTileRows - 1u;
    const bool edge = lane == 63u && xr;
    const unsigned evb = a.eV.beg[q], eve = a.eV.beg[q + 4];
    const unsigned eib = a.eI.beg[q], eie = a.eI.beg[q + 4];
    const unsigned eub = a.eV.beg[q + du], eue = yupd ? a.eV.beg[q + du + 4] : eub;
    // cell x0+4: the next quad's list (ordered by plane), entries of x0+4 only
    const unsigned erb = a.eV.beg[q + dr], ere = xr ? a.eV.beg[q + dr + 4] : erb;
    // Mur faces on x = 0 and x = Nx-1: the quad holding the face runs their
    // pre and post hooks (Engine_Ext_Mur_ABC::DoPreVoltageUpdates /
    // DoPostVoltageUpdates, engine_ext_mur_abc.cpp:68-142) and uses the value
    // the apply hook will set for the face's y and z components in its own
    // current update; it stores that value too, except on the face's rim,
    // where the other faces' hooks still read or write, and which
    // k_mur_apply_rim sets afterwards
    const bool flo = a.mf[0].on && x0 == 0 && (unsigned)ts >= a.mf[0].start_ts;
    const bool fhi = a.mf[1].on && x0 + 4 == Nx && (unsigned)ts >= a.mf[1].start_ts;
    SrcCursor cv = src_cursor(a.eV, evb, eve, z0), ci = src_cursor(a.eI, eib, eie, z0);
    SrcCursor cu = src_cursor(a.eV, eub, eue, z0);
    float4 vm[3] = {}, vu2m = {}, vu0m = {};
    float vr2m = 0.0f, vr1m = 0.0f;
    // I^n of the own quad at plane z-1, carried from the previous plane
    float4 ixp = {}, iyp = {}, izp = {};
    // and its packed table entries (the engine's two half steps share one
    // index array; read again otherwise)
    uint2 wp = make_uint2(0u, 0u);
    const bool same_idx = a.kI.idx == a.kV.idx;
    if (z0) {
        const unsigned long long p0 = (unsigned long long)(z0 - 1) * plane + q;
        ixp = *reinterpret_cast<const float4*>(I0 + p0);
        iyp = *reinterpret_cast<const float4*>(I0 + cs + p0);
        izp = *reinterpret_cast<const float4*>(I0 + 2 * cs + p0);
    }
    for (unsigned z = z0; z <= zlast; ++z) {
        const unsigned long long p = (unsigned long long)z * plane + q;
        const unsigned long long pz = z ? p - plane : p;
        const unsigned b = z % 3u;
        // Mur coefficients of this plane's face points, issued with the field loads
        float mcp = 0.0f, mcpp = 0.0f, mcpp1 = 0.0f;
        if (flo) {
            mcp = a.mf[0].cP[z * a.ny + y];
            mcpp = a.mf[0].cPP[z * a.ny + y];
            if (top && yupd && z < zlast) mcpp1 = a.mf[0].cPP[z * a.ny + y + 1];
        } else if (fhi) {
            mcp = a.mf[1].cP[z * a.ny + y];
            mcpp = a.mf[1].cPP[z * a.ny + y];
        }
        // ---- own V^{n+1}(x0..x0+3, y, z)
        const float4 ix = *reinterpret_cast<const float4*>(I0 + p);
        const float4 iy = *reinterpret_cast<const float4*>(I0 + cs + p);
        const float4 iz = *reinterpret_cast<const float4*>(I0 + 2 * cs + p);
        // x-1 neighbours of element 0: element 3 of the previous lane's quad;
        // lane 0 reads them (or, at x0 = 0, takes the cell itself)
        float iz_l = __shfl_up(iz.w, 1u, 64), iy_l = __shfl_up(iy.w, 1u, 64);
        if (lane == 0u) {
            iz_l = x0 ? I0[2 * cs + p - 1] : iz.x;
            iy_l = x0 ? I0[cs + p - 1] : iy.x;
        }
        float4 v[3];
        unsigned u[4] = {0u, 0u, 0u, 0u};
        uint2 wv = make_uint2(0u, 0u);
        float fy = 0.0f, fz = 0.0f;        // the face element's y and z after the Mur apply
        {
            const float4 iz_ym = *reinterpret_cast<const float4*>(I0 + 2 * cs + p - ym);
            const float4 ix_ym = *reinterpret_cast<const float4*>(I0 + p - ym);
            const float4 ix_m = z ? ixp : ix;
            const float4 iy_m = z ? iyp : iy;
            const float4 iz_xm = make_float4(iz_l, iz.x, iz.y, iz.z);
            const float4 iy_xm = make_float4(iy_l, iy.x, iy.y, iy.z);
            float4 v0[3];
            v0[0] = ld4<NT>(V0 + p);
            v0[1] = ld4<NT>(V0 + cs + p);
            v0[2] = ld4<NT>(V0 + 2 * cs + p);
            if (CC) {
                wv = *reinterpret_cast<const uint2*>(a.kV.idx + p);
                u[0] = wv.x & 0xFFFFu;
                u[1] = wv.x >> 16;
                u[2] = wv.y & 0xFFFFu;
                u[3] = wv.y >> 16;
            }
            float4 c, w;
            quad_coef_n<CC>(a.kV, cs, p, u, 0, c, w);
            v[0] = quad_math(v0[0], c, w, quad_curl(iz, iz_ym, iy, iy_m));
            quad_coef_n<CC>(a.kV, cs, p, u, 1, c, w);
            v[1] = quad_math(v0[1], c, w, quad_curl(ix, ix_m, iz, iz_xm));
            quad_coef_n<CC>(a.kV, cs, p, u, 2, c, w);
            v[2] = quad_math(v0[2], c, w, quad_curl(iy, iy_xm, ix, ix_ym));
            quad_sources_at(a.eV, cv, x0, z, 7u, ts, v);
            // the hook values are kept for k_mur_apply_rim on the face's rim
            const bool keep = active && z < z1 && (y < 2 || y + 2 >= a.ny || z < 2 || z + 2 >= a.zb);
            if (flo) mur_fold_quad<0, 1>(a.mf[0], z * a.ny + y, mcp, mcpp, v0, v, keep, fy, fz);
            else if (fhi) mur_fold_quad<3, 2>(a.mf[1], z * a.ny + y, mcp, mcpp, v0, v, keep, fy, fz);
        }
        // V^{n+1} as the currents read it: the face element after the Mur apply
        float4 vq1 = v[1], vq2 = v[2];
        if (flo) {
            vq1.x = fy;
            vq2.x = fz;
        } else if (fhi) {
            vq1.w = fy;
            vq2.w = fz;
        }
        if (active && z < z1) {
            // face points off the face's rim (no other face's hook reads or
            // writes them) are stored with the apply's value; k_mur_apply_rim
            // sets the rim after the other faces' post hooks
#pragma unroll
            for (int n = 0; n < 3; ++n) st4<NT>(V1 + n * cs + p, v[n]);
            if ((flo || fhi) && y >= 2 && y + 2 < a.ny && z >= 2 && z + 2 < a.zb) {
                const unsigned long long pf = p + (flo ? 0u : 3u);
                V1[cs + pf] = fy;
                V1[2 * cs + pf] = fz;
            }
        }
        s_vz[b][threadIdx.x] = vq2;
        s_vx[b][threadIdx.x] = v[0];
        // element 0 of the next lane's quad at plane z-1 (cell x0+4)
        const float nr2 = __shfl_down(vm[2].x, 1u, 64);
        const float nr1 = __shfl_down(vm[1].x, 1u, 64);
        __syncthreads();
        // ---- I^{n+1}(x0..x0+3, y, z-1) from the carried plane and v
        if (z > z0 && z < a.zb && imask) {
            const unsigned long long pm = pz;
            unsigned um[4] = {0u, 0u, 0u, 0u};
            if (CC) {
                const uint2 wi = same_idx ? wp : *reinterpret_cast<const uint2*>(a.kI.idx + pm);
                um[0] = wi.x & 0xFFFFu;
                um[1] = wi.x >> 16;
                um[2] = wi.y & 0xFFFFu;
                um[3] = wi.y >> 16;
            }
            float4 u2 = vu2m, u0 = vu0m;
            if (!top) {
                const unsigned bm = (z + 2u) % 3u;     // plane z-1
                u2 = s_vz[bm][threadIdx.x + 64u];
                u0 = s_vx[bm][threadIdx.x + 64u];
            }
            const float r2 = lane == 63u ? vr2m : nr2, r1 = lane == 63u ? vr1m : nr1;
            float4 io[3];
            const float4 vzr = make_float4(vm[2].y, vm[2].z, vm[2].w, r2);
            const float4 vyr = make_float4(vm[1].y, vm[1].z, vm[1].w, r1);
            float4 c, w;
            quad_coef_n<CC>(a.kI, cs, pm, um, 0, c, w);
            io[0] = quad_math(ixp, c, w, quad_curl(vm[2], u2, vm[1], vq1));
            quad_coef_n<CC>(a.kI, cs, pm, um, 1, c, w);
            io[1] = quad_math(iyp, c, w, quad_curl(vm[0], v[0], vm[2], vzr));
            quad_coef_n<CC>(a.kI, cs, pm, um, 2, c, w);
            io[2] = quad_math(izp, c, w, quad_curl(vm[1], vyr, vm[0], u0));
            quad_sources_at(a.eI, ci, x0, z - 1, 7u, ts, io);
            if (imask == 15u) {
#pragma unroll
                for (int n = 0; n < 3; ++n) st4<NT>(I1 + n * cs + pm, io[n]);
            } else {
#pragma unroll
                for (int n = 0; n < 3; ++n)
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        if (imask >> j & 1u) I1[n * cs + pm + j] = f4g(io[n], j);
            }
        }
        vm[0] = v[0];
        vm[1] = vq1;
        vm[2] = vq2;
        ixp = ix;
        iyp = iy;
        izp = iz;
        wp = wv;
        if (z < zlast && top && yupd) {
            // row y+1 at plane z (z and x components), outside the tile
            const unsigned long long pu = p + du;
            const float4 ixu = *reinterpret_cast<const float4*>(I0 + pu);
            const float4 iyu = *reinterpret_cast<const float4*>(I0 + cs + pu);
            const float4 izu = *reinterpret_cast<const float4*>(I0 + 2 * cs + pu);
            const float4 iyu_m = *reinterpret_cast<const float4*>(I0 + cs + pz + du);
            float4 iyu_xm;
            if (x0) iyu_xm = ld4u(I0 + cs + pu - 1);
            else iyu_xm = make_float4(iyu.x, iyu.x, iyu.y, iyu.z);
            const float4 u0z = *reinterpret_cast<const float4*>(V0 + 2 * cs + pu);
            const float4 u0x = *reinterpret_cast<const float4*>(V0 + pu);
            unsigned uu[4] = {0u, 0u, 0u, 0u};
            if (CC) quad_idx(a.kV.idx, pu, uu);
            float4 nu[3];
            float4 c, w;
            quad_coef_n<CC>(a.kV, cs, pu, uu, 2, c, w);
            nu[2] = quad_math(u0z, c, w, quad_curl(iyu, iyu_xm, ixu, ix));
            quad_coef_n<CC>(a.kV, cs, pu, uu, 0, c, w);
            nu[0] = quad_math(u0x, c, w, quad_curl(izu, iz, iyu, iyu_m));
            nu[1] = make_float4(0.f, 0.f, 0.f, 0.f);
            quad_sources_at(a.eV, cu, x0, z, 5u, ts, nu);
            if (flo) {
                // the Mur value of row y+1's face cell (z component)
                float mz1 = u0z.y - mcpp1 * u0z.x;
                mz1 = mz1 + mcpp1 * nu[2].y;
                nu[2].x = mz1;
            }
            vu2m = nu[2];
            vu0m = nu[0];
        }
        if (z < zlast && edge) {
            // cell (x0+4, y) at plane z (z and y components), outside the tile
            const unsigned long long pr = p + dr;
            const float ixr = I0[pr], iyr = I0[cs + pr], izr = I0[2 * cs + pr], ixr_ym = I0[pr - ym];
            const float ixr_m = I0[pz + dr];
            const float r0z = V0[2 * cs + pr], r0y = V0[cs + pr];
            const unsigned ur = CC ? a.kV.idx[pr] : 0u;
            const float2 c2 = coef_n<CC>(a.kV, cs, pr, ur, 2), c1 = coef_n<CC>(a.kV, cs, pr, ur, 1);
            const float c2r = iyr - iy.w - ixr + ixr_ym;
            const float c1r = ixr - ixr_m - izr + iz.w;
            float nrz = r0z * c2.x;
            nrz = nrz + c2.y * c2r;
            float nry = r0y * c1.x;
            nry = nry + c1.y * c1r;
            if (erb != ere) {
                nrz = cell_sources(a.eV, erb, ere, x0 + 4, z, 2, ts, nrz);
                nry = cell_sources(a.eV, erb, ere, x0 + 4, z, 1, ts, nry);
            }
            vr2m = nrz;
            vr1m = nry;
        }
    }
}

// ---------------------------------------------------------------------------
// ++numTS after a single-sweep step (Engine::IterateTS, FDTD/engine.cpp:283).
// Every workgroup of the step reads the counter, so the step itself cannot
// advance it; a one-thread launch after it does, with no fence or atomic
// ordering inside the step.
// ---------------------------------------------------------------------------
__global__ void k_tick(unsigned* __restrict__ numts)
{
    if (threadIdx.x == 0) numts[0] = numts[0] + 1u;
}

// ---------------------------------------------------------------------------
// soft / hard source application (Engine_Ext_Excitation::Apply2Voltages /
// Apply2Current, engine_ext_excitation.cpp:33-99).  One thread per (component,
// cell) group adds that cell's entries in list order, as the reference's
// sequential loop does.  numTS is read from the device counter; the
// current-side source runs after the counter advanced, hence ts_bias = -1.
// ---------------------------------------------------------------------------
__global__ void k_excite(float* __restrict__ F, unsigned G, const unsigned* __restrict__ gbeg,
                         const unsigned long long* __restrict__ glin, const float* __restrict__ amp,
                         const unsigned* __restrict__ delay, const float* __restrict__ sig, unsigned length,
                         int period, const unsigned* __restrict__ numts, int ts_bias)
{
    const unsigned g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= G) return;
    const int ts = (int)*numts + ts_bias;
    const unsigned long long i = glin[g];
    float v = F[i];
    for (unsigned k = gbeg[g]; k < gbeg[g + 1]; ++k) v = v + amp[k] * sig[exc_sample(ts, delay[k], period, length)];
    F[i] = v;
}

// ---------------------------------------------------------------------------
// Mur ABC (Engine_Ext_Mur_ABC, engine_ext_mur_abc.cpp:68-178)
// ---------------------------------------------------------------------------
struct MurDev {
    unsigned na, nb;                      // points along nyP, nyPP
    unsigned long long offP, offPP;       // component offsets nyP*cs, nyPP*cs
    const unsigned* face;                 // [na*nb] linear index of the face point
    const unsigned* shift;                // [na*nb] linear index of the neighbour point
    const float* cP;
    const float* cPP;
    float* mP;
    float* mPP;
    unsigned start_ts;
};

__global__ void k_mur_pre(float* __restrict__ V, const MurDev m, const unsigned* __restrict__ numts)
{
    const unsigned k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m.na * m.nb || *numts < m.start_ts) return;
    const unsigned f = m.face[k], s = m.shift[k];
    m.mP[k] = V[m.offP + s] - m.cP[k] * V[m.offP + f];
    m.mPP[k] = V[m.offPP + s] - m.cPP[k] * V[m.offPP + f];
}

__global__ void k_mur_post(const float* __restrict__ V, const MurDev m, const unsigned* __restrict__ numts)
{
    const unsigned k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m.na * m.nb || *numts < m.start_ts) return;
    const unsigned s = m.shift[k];
    m.mP[k] = m.mP[k] + m.cP[k] * V[m.offP + s];
    m.mPP[k] = m.mPP[k] + m.cPP[k] * V[m.offPP + s];
}

__global__ void k_mur_apply(float* __restrict__ V, const MurDev m, const unsigned* __restrict__ numts)
{
    const unsigned k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m.na * m.nb || *numts < m.start_ts) return;
    const unsigned f = m.face[k];
    V[m.offP + f] = m.mP[k];
    V[m.offPP + f] = m.mPP[k];
}

// k_mur_apply at the listed points only (the rim of a face whose other
// points the tiled quad sweep sets)
__global__ void k_mur_apply_rim(float* __restrict__ V, const MurDev m, const unsigned* __restrict__ rim, unsigned n,
                                const unsigned* __restrict__ numts)
{
    const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n || *numts < m.start_ts) return;
    const unsigned k = rim[t];
    const unsigned f = m.face[k];
    V[m.offP + f] = m.mP[k];
    V[m.offPP + f] = m.mPP[k];
}

// The current update (curr_tile's arithmetic, FDTD/engine.cpp:170-222) redone
// from the final voltages V1 and the currents of state n, into I1, for the
// quad sweep's currents that read voltages set after the sweep: the planes
// lo..hi along `axis` next to a Mur face (k_curr_band), and the cells next to
// a deferred voltage source (k_curr_cells).
__device__ __forceinline__ void curr_cell(const GridDesc& g, const float* __restrict__ V1, const float* __restrict__ I0,
                                          float* __restrict__ I1, const CoefArgs& k, unsigned long long p)
{
    const unsigned long long cs = g.cs;
    const unsigned long long dx = 1, dy = g.Nx, dz = g.plane;
    const float* Vx = V1;
    const float* Vy = V1 + cs;
    const float* Vz = V1 + 2 * cs;
    float curl[3];
    curl[0] = Vz[p] - Vz[p + dy] - Vy[p] + Vy[p + dz];
    curl[1] = Vx[p] - Vx[p + dz] - Vz[p] + Vz[p + dx];
    curl[2] = Vy[p] - Vy[p + dx] - Vx[p] + Vx[p + dy];
    float a[3], b[3];
    if (k.idx) {
        const unsigned u = k.idx[p];
        const float4 t0 = k.tab[2 * u], t1 = k.tab[2 * u + 1];
        a[0] = t0.x; b[0] = t0.y;
        a[1] = t0.z; b[1] = t0.w;
        a[2] = t1.x; b[2] = t1.y;
    } else {
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            a[m] = k.cc[m * cs + p];
            b[m] = k.cv[m * cs + p];
        }
    }
#pragma unroll
    for (int m = 0; m < 3; ++m) {
        float f = I0[m * cs + p] * a[m];
        f = f + b[m] * curl[m];
        I1[m * cs + p] = f;
    }
}

// the same at listed cells (each with x, y, z below the last line)
__global__ void k_curr_cells(const GridDesc g, const float* __restrict__ V1, const float* __restrict__ I0,
                             float* __restrict__ I1, const CoefArgs k, const unsigned long long* __restrict__ cells,
                             unsigned n)
{
    const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) curr_cell(g, V1, I0, I1, k, cells[t]);
}

__global__ void k_curr_band(const GridDesc g, const float* __restrict__ V1, const float* __restrict__ I0,
                            float* __restrict__ I1, const CoefArgs k, int axis, unsigned lo, unsigned hi)
{
    const unsigned n[3] = {axis == 0 ? hi - lo + 1 : g.Nx, axis == 1 ? hi - lo + 1 : g.Ny,
                           axis == 2 ? hi - lo + 1 : g.Nz};
    const unsigned long long cnt = (unsigned long long)n[0] * n[1] * n[2];
    const unsigned long long t = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cnt) return;
    unsigned c[3];
    c[0] = (unsigned)(t % n[0]);
    c[1] = (unsigned)((t / n[0]) % n[1]);
    c[2] = (unsigned)(t / ((unsigned long long)n[0] * n[1]));
    c[axis] += lo;
    const unsigned x = c[0], y = c[1], z = c[2];
    if (x + 1 >= g.Nx || y + 1 >= g.Ny || z + 1 >= g.Nz) return;
    curr_cell(g, V1, I0, I1, k, x + (unsigned long long)g.Nx * y + (unsigned long long)g.plane * z);
}

// ---------------------------------------------------------------------------
// lumped RLC (Engine_Ext_LumpedRLC, engine_ext_lumpedRLC.cpp:83-147).  The
// reference rotates its three history buffers once per step (Vd in the pre
// hook, J in the apply hook); after k+1 rotations logical slot i lives in
// buffer (2(k+1) + i) mod 3, so the rotation is derived from numTS.
// ---------------------------------------------------------------------------
struct RlcDev {
    unsigned count;
    const unsigned long long* lin;  // component offset + linear index
    const float* c[9];              // i2v, ilv, vvd, vv2, vj1, vj2, ib0, b1, b2
    float* vd;                      // [3][count]
    float* jn;                      // [3][count]
    float* il;                      // [count]
};

__device__ __forceinline__ unsigned rlc_slot(unsigned ts, unsigned i) { return (2u * ((ts + 1u) % 3u) + i) % 3u; }

__global__ void k_rlc_pre(const RlcDev r, const unsigned* __restrict__ numts)
{
    const unsigned k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= r.count) return;
    const unsigned ts = *numts;
    const float* Vd1 = r.vd + rlc_slot(ts, 1) * r.count;
    r.il[k] = r.il[k] + (r.c[0][k] * r.c[1][k]) * Vd1[k];
}

__global__ void k_rlc_apply(float* __restrict__ V, const RlcDev r, const unsigned* __restrict__ numts)
{
    const unsigned k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= r.count) return;
    const unsigned ts = *numts;
    float* V0 = r.vd + rlc_slot(ts, 0) * r.count;
    const float* V2 = r.vd + rlc_slot(ts, 2) * r.count;
    float* J0 = r.jn + rlc_slot(ts, 0) * r.count;
    const float* J1 = r.jn + rlc_slot(ts, 1) * r.count;
    const float* J2 = r.jn + rlc_slot(ts, 2) * r.count;
    float v = V[r.lin[k]];
    v = r.c[2][k] * (v - r.il[k] + r.c[3][k] * V2[k] + r.c[4][k] * J1[k] + r.c[5][k] * J2[k]);
    V0[k] = v;
    J0[k] = r.c[6][k] * (v - V2[k]) - (r.c[7][k] * r.c[6][k]) * J1[k] - (r.c[8][k] * r.c[6][k]) * J2[k];
    V[r.lin[k]] = v;
}

// ---------------------------------------------------------------------------
// Lorentz / Drude ADE (engine_ext_lorentzmaterial.cpp:79-173,
// engine_ext_dispersive.cpp:76-132)
// ---------------------------------------------------------------------------
struct LorDev {
    unsigned count;
    const unsigned* lin;   // [count]
    const float* cint;     // [3][count]
    const float* cext;
    const float* clor;     // nullptr: Drude/Debye form
    float* ade;            // [3][count]
    float* lor;            // [3][count]
};

__global__ void k_lor_pre(const float* __restrict__ F, unsigned long long cs, const LorDev d)
{
    const unsigned k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= d.count) return;
#pragma unroll
    for (int n = 0; n < 3; ++n) {
        const unsigned j = n * d.count + k;
        const float f = F[n * cs + d.lin[k]];
        if (d.clor) {
            d.lor[j] = d.lor[j] + d.clor[j] * d.ade[j];
            float a = d.ade[j] * d.cint[j];
            d.ade[j] = a + d.cext[j] * (f - d.lor[j]);
        } else {
            float a = d.ade[j] * d.cint[j];
            d.ade[j] = a + d.cext[j] * f;
        }
    }
}

__global__ void k_lor_apply(float* __restrict__ F, unsigned long long cs, const LorDev d)
{
    const unsigned k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= d.count) return;
#pragma unroll
    for (int n = 0; n < 3; ++n) {
        const unsigned long long i = n * cs + d.lin[k];
        F[i] = F[i] - d.ade[n * d.count + k];
    }
}

// ---------------------------------------------------------------------------
// probes: one block per probe; the terms are gathered in parallel and summed
// by one lane in the reference's order so the result is bit-identical
// (double sum for voltages, float sum for currents).
// ---------------------------------------------------------------------------
struct ProbeDev {
    int type;
    unsigned nelem;
    unsigned off;          // offset into the element tables
    unsigned interval, fd_interval, start_ts, stop_ts, capacity;
    unsigned* count;       // [1]
    unsigned* ts_out;      // [capacity]
    double* val_out;       // [capacity*3]
};

constexpr unsigned kProbeChunk = 2048;

// one probe sample by one workgroup (values gathered in parallel, summed by
// one lane in the reference's term order)
__device__ void probe_block(const ProbeDev P, const float* __restrict__ V, const float* __restrict__ I,
                            const unsigned long long* __restrict__ elin, const signed char* __restrict__ esign,
                            const double* __restrict__ ediv, unsigned ts)
{
    __shared__ float vals[kProbeChunk];
    if (ts < P.start_ts || ts > P.stop_ts) return;
    const bool due = (P.interval && ts % P.interval == 0) || (P.fd_interval && ts % P.fd_interval == 0);
    if (!due) return;
    const float* F = (P.type == EMS_PROBE_CURRENT || P.type == EMS_PROBE_HFIELD) ? I : V;
    double acc_d = 0.0;
    float acc_f = 0.0f;
    double out[3] = {0.0, 0.0, 0.0};
    const bool integral = P.type == EMS_PROBE_VOLTAGE || P.type == EMS_PROBE_CURRENT;
    for (unsigned base = 0; base < P.nelem; base += kProbeChunk) {
        const unsigned m = min(kProbeChunk, P.nelem - base);
        // gathered with the term's sign (a - b and a + (-b) round alike), so
        // the ordered sum below reads nothing from global memory
        for (unsigned e = threadIdx.x; e < m; e += blockDim.x) {
            const unsigned j = P.off + base + e;
            const float v = F[elin[j]];
            vals[e] = integral && esign[j] <= 0 ? -v : v;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (unsigned e = 0; e < m; ++e) {
                const unsigned j = P.off + base + e;
                if (P.type == EMS_PROBE_VOLTAGE) {
                    acc_d = acc_d + (double)vals[e];
                } else if (P.type == EMS_PROBE_CURRENT) {
                    acc_f = acc_f + vals[e];
                } else if (base + e < 3) {
                    const double d = ediv[j];
                    const double r = d != 0.0 ? (double)vals[e] / d : 0.0;
                    if (base + e == 0) out[0] = r;
                    if (base + e == 1) out[1] = r;
                    if (base + e == 2) out[2] = r;
                }
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const unsigned slot = *P.count;
        if (slot < P.capacity) {
            // the three results as plain selects, stored unconditionally
            const double r0 = P.type == EMS_PROBE_VOLTAGE ? acc_d : P.type == EMS_PROBE_CURRENT ? (double)acc_f : out[0];
            const double r1 = integral ? 0.0 : out[1];
            const double r2 = integral ? 0.0 : out[2];
            P.ts_out[slot] = ts;
            P.val_out[3 * slot] = r0;
            P.val_out[3 * slot + 1] = r1;
            P.val_out[3 * slot + 2] = r2;
            *P.count = slot + 1;
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_probe(const float* __restrict__ V, const float* __restrict__ I,
                                                  const ProbeDev* __restrict__ probes,
                                                  const unsigned long long* __restrict__ elin,
                                                  const signed char* __restrict__ esign,
                                                  const double* __restrict__ ediv,
                                                  const unsigned* __restrict__ numts)
{
    probe_block(probes[blockIdx.x], V, I, elin, esign, ediv, *numts);
}

// ---------------------------------------------------------------------------
// Batched engine: many independent instances (own mesh, operator, UPML,
// sources, probes) advance together.  Each half step is one persistent
// launch: a workgroup walks the work table (instance, tile) with a stride of
// the grid size; the per-tile arithmetic is that of k_volt / k_curr, so every
// instance is bit-identical to its own engine.
// ---------------------------------------------------------------------------
struct BatchInst {
    GridDesc g;
    float* V;
    float* I;
    CoefArgs kV, kI;
    PmlArgs pV, pI;
    unsigned* numts;
    const float* sig[2];
    unsigned sigLen;
    int period;
    ExcArgs vsrc;                  // voltage sources added by the voltage half step (beg null: by k_batch_excite)
};

template <int PM, bool CC, bool CURR>
__global__ __launch_bounds__(kBlock) void k_batch_half(const BatchInst* __restrict__ inst,
                                                       const uint2* __restrict__ items, unsigned nitems)
{
    for (unsigned w = xcd_tile(blockIdx.x, gridDim.x); w < nitems; w += gridDim.x) {
        const uint2 it = items[w];   // (instance, tile)
        const BatchInst& B = inst[it.x];
        if (CURR) {
            if (it.y == 0 && threadIdx.x == 0) atomicAdd(B.numts, 1u);
            curr_tile<PM, CC, false>(B.g, it.y, B.I, B.V, B.kI, B.pI);
        } else {
            volt_tile<PM, CC, false>(B.g, it.y, B.V, B.I, B.kV, B.pV, B.vsrc, B.numts);
        }
    }
}

// sources of all instances: one thread per (instance, component, cell) group
struct BatchExcGroup {
    unsigned inst;
    unsigned beg, end;             // entries in amp / delay
    unsigned long long lin;        // component * cs + cell
};

__global__ void k_batch_excite(const BatchInst* __restrict__ inst, const BatchExcGroup* __restrict__ grp, unsigned G,
                               const float* __restrict__ amp, const unsigned* __restrict__ delay, int curr, int ts_bias)
{
    const unsigned g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= G) return;
    const BatchExcGroup E = grp[g];
    const BatchInst& B = inst[E.inst];
    float* F = curr ? B.I : B.V;
    const float* sig = B.sig[curr];
    const int ts = (int)*B.numts + ts_bias;
    float v = F[E.lin];
    for (unsigned k = E.beg; k < E.end; ++k) v = v + amp[k] * sig[exc_sample(ts, delay[k], B.period, B.sigLen)];
    F[E.lin] = v;
}

// probes of all instances, one workgroup each; a probe not due returns at once
struct BatchProbe {
    ProbeDev P;
    unsigned inst;
    const unsigned long long* elin;
    const signed char* esign;
    const double* ediv;
};

__global__ __launch_bounds__(kBlock) void k_batch_probe(const BatchInst* __restrict__ inst,
                                                        const BatchProbe* __restrict__ probes)
{
    const BatchProbe& Q = probes[blockIdx.x];
    const BatchInst& B = inst[Q.inst];
    probe_block(Q.P, B.V, B.I, Q.elin, Q.esign, Q.ediv, *B.numts);
}

// ---------------------------------------------------------------------------
// field dumps: ProcessFieldsFD::Process (processfields_fd.cpp:72-107, running
// complex<float> sums) and ProcessFieldsTD::Process (processfields_td.cpp:50-91,
// stored samples) of the interpolated E or H field
// (Engine_Interface_FDTD, engine_interface_fdtd.cpp:63-204)
// ---------------------------------------------------------------------------
// A field dump: frequency domain (td = 0; running DFT of the samples) or
// time domain (td = 1; every sample stored).  Samples are taken after the
// steps ts = first + k * fd_interval, k < nsamples.
struct DumpDev {
    unsigned s[3], n[3];          // box start and extent (mesh indices)
    unsigned N[3];
    unsigned fd_interval, nfreq, nsamples;
    unsigned first;               // step of the first sample (0 for frequency-domain dumps)
    int td;
    int field;                    // 0: E from the voltages, 1: H from the currents
    int interp;                   // 0: none, 1: node, 2: cell interpolation
    const float* factors;         // [nsamples][nfreq][2] (frequency domain)
    const double* el[3];          // edge lengths per axis: primary (E) or dual (H)
    float* acc;                   // FD: [nfreq][3][npts][2]; TD: [nsamples][3][npts]
};

__host__ __device__ inline bool dump_sample(const DumpDev& d, unsigned ts, unsigned* k)
{
    if (ts < d.first || (ts - d.first) % d.fd_interval) return false;
    *k = (ts - d.first) / d.fd_interval;
    return *k < d.nsamples;
}

// Engine_Interface_FDTD::GetRawInterpolatedField (E, engine_interface_fdtd.cpp:63-124)
// and GetRawInterpolatedDualField (H, :148-204), component n at node pos, in
// double; without interpolation GetRawField / GetRawDualField (:263-268,
// :134-146), the stored value over its edge length, zero for a zero length.  raw(dx, dy, dz) is component n's stored value at pos + (dx, dy, dz)
// (offsets -1, 0, +1); el: the component's edge lengths (primary for E, dual
// for H).  Shared by the dump kernel and the host-side point read.
template <class Raw>
__host__ __device__ inline double interp_value(int field, int interp, const unsigned* N, const double* el,
                                               const unsigned* pos, int n, Raw raw)
{
    const int nP = (n + 1) % 3, nPP = (n + 2) % 3;
    // the value at pos + sa*e_a + sb*e_b over delta (zero for a zero length)
    auto rv = [&](int a, int sa, int b, int sb, double delta) -> double {
        int o[3] = {0, 0, 0};
        if (a >= 0) o[a] += sa;
        if (b >= 0) o[b] += sb;
        return delta ? (double)raw(o[0], o[1], o[2]) / delta : 0.0;
    };
    const unsigned pn = pos[n];
    if (interp == 0) return rv(-1, 0, -1, 0, el[pn]);
    if (field == 0) {
        if (interp == 1) {
            if (pn == N[n] - 1) return rv(n, -1, -1, 0, el[pn - 1]);
            const double delta = el[pn];
            double out = rv(-1, 0, -1, 0, delta);
            if (delta == 0.0) return 0.0;
            if (pn == 0) return out;
            const double deltaDown = el[pn - 1];
            const double deltaRel = delta / (delta + deltaDown);
            return out * (1.0 - deltaRel) + rv(n, -1, -1, 0, deltaDown) * deltaRel;
        }
        if (pos[0] == N[0] - 1 || pos[1] == N[1] - 1 || pos[2] == N[2] - 1) return 0.0;
        const double delta = el[pn];
        double out = rv(-1, 0, -1, 0, delta);
        out += rv(nP, 1, -1, 0, delta);
        out += rv(nP, 1, nPP, 1, delta);
        out += rv(nPP, 1, -1, 0, delta);
        return out / 4;
    }
    if (interp == 1) {
        if (pos[0] == N[0] - 1 || pos[1] == N[1] - 1 || pos[2] == N[2] - 1 || pos[nP] == 0 || pos[nPP] == 0)
            return 0.0;
        const double delta = el[pn];
        double out = rv(-1, 0, -1, 0, delta);
        out += rv(nP, -1, -1, 0, delta);
        out += rv(nP, -1, nPP, -1, delta);
        out += rv(nPP, -1, -1, 0, delta);
        return out / 4;
    }
    if (pn >= N[n] - 1) return 0.0;
    const double delta = el[pn];
    const double deltaUp = el[pn + 1];
    const double deltaRel = delta / (delta + deltaUp);
    return rv(-1, 0, -1, 0, delta) * (1.0 - deltaRel) + rv(n, 1, -1, 0, deltaUp) * deltaRel;
}

__device__ inline double dump_value(const DumpDev& d, const float* F, unsigned long long cs, const unsigned* pos,
                                    unsigned long long base, const unsigned long long* stride, int n)
{
    const float* C = F + (unsigned long long)n * cs;
    auto raw = [&](int dx, int dy, int dz) -> float {
        long long off = (long long)base;
        off += (long long)dx * (long long)stride[0];
        off += (long long)dy * (long long)stride[1];
        off += (long long)dz * (long long)stride[2];
        return C[off];
    };
    return interp_value(d.field, d.interp, d.N, d.el[n], pos, n, raw);
}

// ProcessFieldsFD::Process (processfields_fd.cpp:72-107): every fd_interval
// steps, field (float) times exp(-2 pi i f T) * 2 dT fd_interval (complex<float>)
// added to the running sum of each frequency.  One thread per dump point.
__global__ void k_fd_dump(const float* __restrict__ F, unsigned long long cs, unsigned Nx, unsigned plane, const DumpDev d,
                          const unsigned* __restrict__ numts)
{
    const unsigned ts = *numts;
    unsigned ks;
    if (!dump_sample(d, ts, &ks)) return;
    const unsigned npts = d.n[0] * d.n[1] * d.n[2];
    const unsigned k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= npts) return;
    // dump-box N-I-J-K order: z fastest
    const unsigned lz = k % d.n[2];
    const unsigned ly = (k / d.n[2]) % d.n[1];
    const unsigned lx = k / (d.n[2] * d.n[1]);
    const unsigned pos[3] = {d.s[0] + lx, d.s[1] + ly, d.s[2] + lz};
    const unsigned long long base = pos[0] + (unsigned long long)Nx * pos[1] + (unsigned long long)plane * pos[2];
    const unsigned long long stride[3] = {1ull, (unsigned long long)Nx, (unsigned long long)plane};
    float E[3];
#pragma unroll
    for (int n = 0; n < 3; ++n) E[n] = (float)dump_value(d, F, cs, pos, base, stride, n);
    if (d.td) {
        // ProcessFieldsTD::Process (processfields_td.cpp:50-91): the sample itself
#pragma unroll
        for (int n = 0; n < 3; ++n) d.acc[((unsigned long long)ks * 3 + n) * npts + k] = E[n];
        return;
    }
    for (unsigned f = 0; f < d.nfreq; ++f) {
        const float er = d.factors[(ks * d.nfreq + f) * 2];
        const float ei = d.factors[(ks * d.nfreq + f) * 2 + 1];
#pragma unroll
        for (int n = 0; n < 3; ++n) {
            float* a = d.acc + (((unsigned long long)f * 3 + n) * npts + k) * 2;
            a[0] = a[0] + E[n] * er;
            a[1] = a[1] + E[n] * ei;
        }
    }
}

// ---------------------------------------------------------------------------
// energy (Engine_Interface_FDTD::CalcFastEnergy): per-block double partials
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_energy(const float* __restrict__ V, const float* __restrict__ I,
                                                   unsigned long long cs, unsigned Nx, unsigned Ny, unsigned Nz,
                                                   double* __restrict__ part)
{
    __shared__ double se[kBlock], sh[kBlock];
    const unsigned long long plane = (unsigned long long)Nx * Ny;
    const unsigned long long tot = plane * Nz;
    double e = 0.0, h = 0.0;
    for (unsigned long long p = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x; p < tot;
         p += (unsigned long long)gridDim.x * blockDim.x) {
        const unsigned long long z = p / plane;
        const unsigned long long r = p - z * plane;
        const unsigned long long y = r / Nx, x = r - y * Nx;
        if (x + 1 >= Nx || y + 1 >= Ny || z + 1 >= Nz) continue;
        for (int n = 0; n < 3; ++n) {
            const double v = V[n * cs + p], c = I[n * cs + p];
            e += v * v;
            h += c * c;
        }
    }
    se[threadIdx.x] = e;
    sh[threadIdx.x] = h;
    __syncthreads();
    for (unsigned s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            se[threadIdx.x] += se[threadIdx.x + s];
            sh[threadIdx.x] += sh[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = se[0];
        part[2 * blockIdx.x + 1] = sh[0];
    }
}

// one thread: the energy from k_energy's partial sums, summed in workgroup
// order (as ems_energy sums them on the host)
__global__ void k_energy_total(const double* __restrict__ part, unsigned nb, double* __restrict__ out)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double e = 0.0, m = 0.0;
    for (unsigned b = 0; b < nb; ++b) {
        e += part[2 * b];
        m += part[2 * b + 1];
    }
    *out = 8.85418781762e-12 * e + 1.256637062e-6 * m;
}

}  // namespace

// Workgroups of the energy reduction, and the device scratch of the
// allocation-free energy entry point (ems_energy_noalloc): the partial sums
// of those workgroups, two doubles each (16 KiB; every device the library runs
// on holds one copy until the process ends).  Host code never passes the
// array's name to a kernel or a copy: it uses the per-device address from
// hipGetSymbolAddress (scratch_setup).
constexpr unsigned kEnergyBlocks = 1024;
static __device__ __align__(256) double scratch[2 * kEnergyBlocks];

namespace {

// ---------------------------------------------------------------------------
// Adjoint of the timestep (discrete adjoint of the two half steps with UPML),
// for gradients of a DFT-of-probe objective with respect to the coefficients.
// Forward half step of one cell (n = x, y, z), curl c, UPML flux F:
//   plain: X' = cc*X + cv*c
//   UPML : f = cc*F + cv*c;  X' = c0*X - cfo*F + cfn*f;  F' = f
// The transposed half step maps the adjoint (lX', lF') of the new state to
// the adjoint of the old state and the adjoint of the curl:
//   plain: lX = cc*lX';  lc = cv*lX'
//   UPML : lf = lF' + cfn*lX';  lX = c0*lX';  lF = -cfo*lX' + cc*lf;  lc = cv*lf
// and lc is scattered back through the transposed curl, evaluated here in
// gather form (terms whose forward index clamps to the cell itself vanish,
// as the forward curl cancels there).
// ---------------------------------------------------------------------------
struct AdjBox {
    unsigned x0, y0, z0, nx, ny, nz;   // inclusive-exclusive box of edges
    unsigned long long n;              // nx*ny*nz
};

__device__ __forceinline__ bool in_adj_box(const AdjBox& b, unsigned x, unsigned y, unsigned z, unsigned long long& k)
{
    if (x < b.x0 || y < b.y0 || z < b.z0 || x >= b.x0 + b.nx || y >= b.y0 + b.ny || z >= b.z0 + b.nz) return false;
    k = (x - b.x0) + (unsigned long long)b.nx * ((y - b.y0) + (unsigned long long)b.ny * (z - b.z0));
    return true;
}

// UPML coefficients c0, cfo, cfn of position li
template <int PM>
__device__ __forceinline__ void pml_coefs(const PmlArgs& P, unsigned long long li, float k0[3], float ko[3], float kn[3])
{
    if (PM == kPmlPlain) {
#pragma unroll
        for (int n = 0; n < 3; ++n) {
            k0[n] = P.coef[n * P.npos + li];
            ko[n] = P.coef[(3 + n) * P.npos + li];
            kn[n] = P.coef[(6 + n) * P.npos + li];
        }
    } else {
        const unsigned u = P.pidx[li];
        const float4 t0 = P.ptab[3 * u], t1 = P.ptab[3 * u + 1], t2 = P.ptab[3 * u + 2];
        k0[0] = t0.x; k0[1] = t0.y; k0[2] = t0.z;
        ko[0] = t0.w; ko[1] = t1.x; ko[2] = t1.y;
        kn[0] = t1.z; kn[1] = t1.w; kn[2] = t2.x;
    }
}

// transposed current half step, one thread per (x, y) column sweeping z.
// lI' = lI + curl_I^T(B) (B: curl adjoint left by the transposed voltage
// half step of the later timestep); writes lI, lFI and A = curl_V adjoint.
template <int PM, bool CC>
__global__ __launch_bounds__(kBlock) void k_adj_curr(const GridDesc g, float* __restrict__ lI,
                                                     const float* __restrict__ B, float* __restrict__ A,
                                                     const CoefArgs k, const PmlArgs pml, float* __restrict__ lFI)
{
    const unsigned t = xcd_tile(blockIdx.x, gridDim.x);
    const unsigned qt = t % g.qtiles, zc = t / g.qtiles;
    unsigned x, y;
    if (!tile_column(g, qt, x, y)) return;
    const unsigned z0 = zc * g.zdepth, z1 = min(z0 + g.zdepth, g.Nz);
    const unsigned long long cs = g.cs;
    const unsigned Nx = g.Nx;
    const bool col_upd = x + 1 < g.Nx && y + 1 < g.Ny;
    for (unsigned z = z0; z < z1; ++z) {
        const unsigned long long p = x + (unsigned long long)Nx * y + (unsigned long long)g.plane * z;
        // lI' = lI + curl_I^T(B); curl_I of the voltage update uses lower
        // neighbours, clamped to the cell itself at index 0
        float l[3];
        {
            const float* B0 = B;
            const float* B1 = B + cs;
            const float* B2 = B + 2 * cs;
            float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f;
            if (z > 0) { s0 += B1[p]; s1 -= B0[p]; }
            if (z + 1 < g.Nz) { s0 -= B1[p + g.plane]; s1 += B0[p + g.plane]; }
            if (y > 0) { s0 -= B2[p]; s2 += B0[p]; }
            if (y + 1 < g.Ny) { s0 += B2[p + Nx]; s2 -= B0[p + Nx]; }
            if (x > 0) { s1 += B2[p]; s2 -= B1[p]; }
            if (x + 1 < g.Nx) { s1 -= B2[p + 1]; s2 += B1[p + 1]; }
            l[0] = lI[p] + s0;
            l[1] = lI[cs + p] + s1;
            l[2] = lI[2 * cs + p] + s2;
        }
        const bool upd = col_upd && z + 1 < g.Nz;
        float lo[3], a[3] = {0.0f, 0.0f, 0.0f};
        unsigned long long li = 0;
        const bool in_pml = PM != kNoPml && pml_at(pml, (int)x, (int)y, (int)z, li);
        if (upd) {
            float ca[3], cb[3];
            load_coef<CC>(k, cs, p, CC ? k.idx[p] : 0u, ca, cb);
            if (in_pml) {
                float k0[3], ko[3], kn[3];
                pml_coefs<PM>(pml, li, k0, ko, kn);
#pragma unroll
                for (int n = 0; n < 3; ++n) {
                    const float lf = lFI[n * pml.npos + li] + kn[n] * l[n];
                    lo[n] = k0[n] * l[n];
                    lFI[n * pml.npos + li] = ca[n] * lf - ko[n] * l[n];
                    a[n] = cb[n] * lf;
                }
            } else {
#pragma unroll
                for (int n = 0; n < 3; ++n) {
                    lo[n] = ca[n] * l[n];
                    a[n] = cb[n] * l[n];
                }
            }
        } else if (in_pml) {
            // UPML hooks alone: X' = c0*X + (cfn - cfo)*F, F' = F
            float k0[3], ko[3], kn[3];
            pml_coefs<PM>(pml, li, k0, ko, kn);
#pragma unroll
            for (int n = 0; n < 3; ++n) {
                lo[n] = k0[n] * l[n];
                lFI[n * pml.npos + li] = lFI[n * pml.npos + li] + (kn[n] - ko[n]) * l[n];
            }
        } else {
#pragma unroll
            for (int n = 0; n < 3; ++n) lo[n] = l[n];
        }
#pragma unroll
        for (int n = 0; n < 3; ++n) {
            lI[n * cs + p] = lo[n];
            A[n * cs + p] = a[n];
        }
    }
}

// transposed voltage half step: lV' = lV + curl_V^T(A); writes lV, lFV,
// B = curl_I adjoint, and accumulates dJ/dvi, dJ/dvv on the box edges from
// the recorded forward curl and voltage of the step.
template <int PM, bool CC>
__global__ __launch_bounds__(kBlock) void k_adj_volt(const GridDesc g, float* __restrict__ lV,
                                                     const float* __restrict__ A, float* __restrict__ B,
                                                     const CoefArgs k, const PmlArgs pml, float* __restrict__ lFV,
                                                     const AdjBox box, const float* __restrict__ hc,
                                                     const float* __restrict__ hv, double* __restrict__ gvi,
                                                     double* __restrict__ gvv)
{
    const unsigned t = xcd_tile(blockIdx.x, gridDim.x);
    const unsigned qt = t % g.qtiles, zc = t / g.qtiles;
    unsigned x, y;
    if (!tile_column(g, qt, x, y)) return;
    const unsigned z0 = zc * g.zdepth, z1 = min(z0 + g.zdepth, g.Nz);
    const unsigned long long cs = g.cs;
    const unsigned Nx = g.Nx;
    for (unsigned z = z0; z < z1; ++z) {
        const unsigned long long p = x + (unsigned long long)Nx * y + (unsigned long long)g.plane * z;
        // lV' = lV + curl_V^T(A); curl_V of the current update uses upper
        // neighbours, A is zero where the current update does not run
        float l[3];
        {
            const float* A0 = A;
            const float* A1 = A + cs;
            const float* A2 = A + 2 * cs;
            float s0 = A1[p] - A2[p];
            float s1 = A2[p] - A0[p];
            float s2 = A0[p] - A1[p];
            if (z > 0) { s0 -= A1[p - g.plane]; s1 += A0[p - g.plane]; }
            if (y > 0) { s0 += A2[p - Nx]; s2 -= A0[p - Nx]; }
            if (x > 0) { s1 -= A2[p - 1]; s2 += A1[p - 1]; }
            l[0] = lV[p] + s0;
            l[1] = lV[cs + p] + s1;
            l[2] = lV[2 * cs + p] + s2;
        }
        float ca[3], cb[3], lo[3], b[3];
        load_coef<CC>(k, cs, p, CC ? k.idx[p] : 0u, ca, cb);
        unsigned long long li = 0;
        const bool in_pml = PM != kNoPml && pml_at(pml, (int)x, (int)y, (int)z, li);
        if (in_pml) {
            float k0[3], ko[3], kn[3];
            pml_coefs<PM>(pml, li, k0, ko, kn);
#pragma unroll
            for (int n = 0; n < 3; ++n) {
                const float lf = lFV[n * pml.npos + li] + kn[n] * l[n];
                lo[n] = k0[n] * l[n];
                lFV[n * pml.npos + li] = ca[n] * lf - ko[n] * l[n];
                b[n] = cb[n] * lf;
            }
        } else {
#pragma unroll
            for (int n = 0; n < 3; ++n) {
                lo[n] = ca[n] * l[n];
                b[n] = cb[n] * l[n];
            }
            unsigned long long kb;
            if (in_adj_box(box, x, y, z, kb)) {
#pragma unroll
                for (int n = 0; n < 3; ++n) {
                    gvi[n * box.n + kb] += (double)l[n] * (double)hc[n * box.n + kb];
                    gvv[n * box.n + kb] += (double)l[n] * (double)hv[n * box.n + kb];
                }
            }
        }
#pragma unroll
        for (int n = 0; n < 3; ++n) {
            lV[n * cs + p] = lo[n];
            B[n * cs + p] = b[n];
        }
    }
}

// forward curl of the currents and the voltages on the box edges before a
// step (the inputs of the coming voltage update)
__global__ void k_adj_record(const GridDesc g, const float* __restrict__ V, const float* __restrict__ I,
                             const AdjBox box, float* __restrict__ hc, float* __restrict__ hv)
{
    const unsigned long long kk = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x;
    if (kk >= box.n) return;
    const unsigned x = box.x0 + (unsigned)(kk % box.nx);
    const unsigned y = box.y0 + (unsigned)((kk / box.nx) % box.ny);
    const unsigned z = box.z0 + (unsigned)(kk / ((unsigned long long)box.nx * box.ny));
    const unsigned long long cs = g.cs;
    const unsigned long long p = x + (unsigned long long)g.Nx * y + (unsigned long long)g.plane * z;
    const unsigned long long xm = x ? 1 : 0, ym = y ? g.Nx : 0, zm = z ? g.plane : 0;
    const float* Ix = I;
    const float* Iy = I + cs;
    const float* Iz = I + 2 * cs;
    hc[kk] = Iz[p] - Iz[p - ym] - Iy[p] + Iy[p - zm];
    hc[box.n + kk] = Ix[p] - Ix[p - zm] - Iz[p] + Iz[p - xm];
    hc[2 * box.n + kk] = Iy[p] - Iy[p - xm] - Ix[p] + Ix[p - ym];
#pragma unroll
    for (int n = 0; n < 3; ++n) hv[n * box.n + kk] = V[n * cs + p];
}

// objective probes: F_p += (V[lin_p] / div_p) * w, w = exp(-2 pi i f t) * dT
// at t = ts*dT (voltages after step ts)
struct AdjProbe {
    unsigned long long lin;
    double div;
};

__global__ void k_adj_dft(const float* __restrict__ V, const AdjProbe* __restrict__ pr, unsigned np, double f,
                          double dT, const unsigned* __restrict__ numts, double* __restrict__ F)
{
    const unsigned p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const double t = (double)*numts * dT;
    const double ph = -2.0 * 3.141592653589793238462643383279 * f * t;
    const double v = (double)V[pr[p].lin] / pr[p].div;
    F[2 * p] += v * cos(ph) * dT;
    F[2 * p + 1] += v * sin(ph) * dT;
}

// adjoint source of J = sum |F_p|^2 at step ts: dJ/dV = 2 Re(conj(F) w) / div
__global__ void k_adj_src(float* __restrict__ lV, const AdjProbe* __restrict__ pr, unsigned np, double f, double dT,
                          unsigned ts, const double* __restrict__ F)
{
    const unsigned p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const double t = (double)ts * dT;
    const double ph = -2.0 * 3.141592653589793238462643383279 * f * t;
    const double g = 2.0 * (F[2 * p] * cos(ph) * dT + F[2 * p + 1] * sin(ph) * dT) / pr[p].div;
    lV[pr[p].lin] = (float)((double)lV[pr[p].lin] + g);
}

// ---------------------------------------------------------------------------
// steady-state detection (Engine_Ext_SteadyState::Apply2Voltages,
// FDTD/extensions/engine_ext_steadystate.cpp:50-107).  Priority +2e6: it runs
// before every other apply hook, on the voltages of the completed voltage
// update.  Every step records the probe voltages at TS % 2p; at TS % p == 0,
// TS >= 2p it compares the last two periods and the energy change.
// ---------------------------------------------------------------------------
struct SteadyDev {
    unsigned n, p, cap;
    const unsigned long long* lin;   // [n] component * cs + cell
    float* rec;                      // [n][2p]
    double* state;                   // [0] last max diff, [1] last total energy, [2] evaluations
    unsigned* hist_ts;               // [cap]
    double* hist_diff;               // [cap]
};

__global__ void k_ss_record(const float* __restrict__ V, const SteadyDev s, const unsigned* __restrict__ numts)
{
    const unsigned k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= s.n) return;
    const unsigned rel = *numts % (2 * s.p);
    s.rec[(size_t)k * 2 * s.p + rel] = V[s.lin[k]];
}

// one thread: the reference's evaluation, energy from k_energy's partials
__global__ void k_ss_eval(const SteadyDev s, const double* __restrict__ part, unsigned nb,
                          const unsigned* __restrict__ numts)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const unsigned TS = *numts, p = s.p;
    const unsigned rel_pos = TS % (2 * p);
    double e = 0, m = 0;
    for (unsigned b = 0; b < nb; ++b) {
        e += part[2 * b];
        m += part[2 * b + 1];
    }
    const double curr_total_energy = 8.85418781762e-12 * e + 1.256637062e-6 * m;
    bool no_valid = true;
    double last_max_diff = 0;
    const double last_total_energy = s.state[1];
    if (last_total_energy > 0) {
        last_max_diff = fabs(curr_total_energy - last_total_energy) / last_total_energy;
        no_valid = false;
    }
    unsigned old_pos = 0, new_pos = p;
    if (rel_pos <= p) {
        new_pos = 0;
        old_pos = p;
    }
    double max_pow = 0;
    for (unsigned n = 0; n < s.n; ++n) {
        const float* buf = s.rec + (size_t)n * 2 * p;
        double cp = 0;
        for (unsigned nt = 0; nt < p; ++nt) cp += (double)buf[nt + new_pos] * (double)buf[nt + new_pos];
        max_pow = fmax(max_pow, cp);
    }
    for (unsigned n = 0; n < s.n; ++n) {
        const float* buf = s.rec + (size_t)n * 2 * p;
        double cp = 0, dp = 0;
        for (unsigned nt = 0; nt < p; ++nt) {
            const double a = buf[nt + new_pos], b = buf[nt + old_pos];
            cp += a * a;
            dp += (b - a) * (b - a);
        }
        if (cp > max_pow * 1e-2) {
            last_max_diff = fmax(last_max_diff, dp / cp);
            no_valid = false;
        }
    }
    if (no_valid || last_max_diff > 1) last_max_diff = 1;
    s.state[0] = last_max_diff;
    s.state[1] = curr_total_energy;
    const unsigned k = (unsigned)s.state[2];
    if (k < s.cap) {
        s.hist_ts[k] = TS;
        s.hist_diff[k] = last_max_diff;
    }
    s.state[2] = (double)(k + 1);
}

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
};

void parallel_rows(size_t n, const std::function<void(size_t, size_t)>& fn);

}  // namespace

// ---------------------------------------------------------------------------
// engine state
// ---------------------------------------------------------------------------
struct ems_engine {
    int device = 0;
    hipStream_t stream = nullptr;
    GridDesc g{};
    unsigned N[3] = {0, 0, 0};
    unsigned long long cs = 0;
    float *V = nullptr, *I = nullptr;          // current state
    float *V_alt = nullptr, *I_alt = nullptr;  // second buffers of the fused step (state n+1)
    float* V_base = nullptr;                   // V at set-up (buffer parity of the fused step)
    // coefficient layout: plain arrays or per-cell index + tables
    int layout_req = EMS_LAYOUT_AUTO;
    bool cmp = false;
    float *vv = nullptr, *vi = nullptr, *ii = nullptr, *iv = nullptr;
    unsigned short* cidx = nullptr;
    float* ctab[2] = {nullptr, nullptr};
    unsigned n_unique = 0;
    CoefArgs kV{}, kI{};
    unsigned* numts_dev = nullptr;
    unsigned numTS = 0;
    bool inited = false;
    bool async = false;
    std::vector<void*> allocs;

    // excitation: entries grouped per (component, cell), each group in list order
    struct Exc {
        unsigned K = 0, G = 0;
        unsigned* gbeg = nullptr;         // [G + 1]
        unsigned long long* glin = nullptr;   // [G] component * cs + cell
        float* amp = nullptr;             // [K] in group order
        unsigned* delay = nullptr;        // [K]
        std::vector<ExcEnt> host;         // list order, for the fused step's tile lists
        std::vector<unsigned> h_gbeg;      // host copies of the groups (batched engine)
        std::vector<unsigned long long> h_glin;
        std::vector<float> h_amp;
        std::vector<unsigned> h_delay;
    } exc[2];
    float* sig[2] = {nullptr, nullptr};
    unsigned sigLen = 0;
    int period = 0;
    bool has_exc = false;

    // UPML: host staging per slab until init
    struct SlabHost {
        SlabTab tab;
        unsigned num[3] = {0, 0, 0};      // slab extent (mesh lines)
        unsigned long long cnt = 0;       // cells of the slab
        std::vector<float> v[3], i[3];   // vv, vvfo, vvfn / ii, iifo, iifn, [3][cnt] x-fastest
    };
    std::vector<SlabHost> slabs;
    PmlArgs pmlV{}, pmlI{};
    float* flux_alt[2] = {nullptr, nullptr};   // second flux buffers of the fused step (V, I)
    // fused single-sweep step
    int fused_req = -1;                   // -1 auto, 0 two half steps, 1 LDS sweep, 2 column sweep, 3 quad sweep
    int fused = 0;                        // active form: 0 two half steps, 1 k_step, 2 k_step2
    bool nt_fields = false;               // non-temporal own-field accesses in the half steps
    int quad_req = -1;                    // -1 auto, 0 one cell per lane, 1 quads (four x cells per lane)
    bool lds_tab = false;                 // quad kernels read the compressed tables from LDS
    bool quad = false;                    // the half steps run k_volt4 / k_curr4
    GridDesc gq{};                        // z depth of the quad half steps
    // quad boxes: the whole mesh, or on a UPML mesh the box without UPML
    // cells (qbox_in, launched without UPML code) and the boxes around it
    QBoxes qbox_all{}, qbox_in{}, qbox_out{};
    unsigned qblk_all = 0, qblk_in = 0, qblk_out = 0;
    QBox* qbox_dev = nullptr;             // the three box lists, kMaxQBoxes each
    // voltage sources added by k_volt4 / k_volt (per-quad or per-column lists;
    // beg null: k_excite adds them)
    ExcArgs qsrc{};
    bool qsplit = false;
    unsigned ntx = 0, nty = 0, fzdepth = 0, fnzc = 0;
    bool q4tile = false;                  // form 3 runs k_step4t (64 x 4 quad tiles)
    unsigned q4ntq = 0;                   // its tiles along x
    int mur_fold[2] = {-1, -1};           // Mur faces (index) on x = 0 / x = Nx-1 whose pre and post hooks k_step4t runs
    unsigned* mur_rim[2] = {nullptr, nullptr};    // their rim points (y or z within 2 lines of the mesh edge)
    unsigned nrim[2] = {0, 0};
    ExcArgs eV{}, eI{};
    int pml_mode = kNoPml;
    unsigned n_unique_pml = 0;
    // Mur faces in creation order
    std::vector<MurDev> mur;
    // Mur face geometry (normal axis, face line, neighbour line), creation order
    struct MurGeo {
        int ny;
        unsigned line, shift;
    };
    std::vector<MurGeo> mur_geo;
    // quad sweep with Mur faces: sources on face / neighbour lines added after
    // the Mur hooks (exc_def[0]), the current update redone on the planes next
    // to each face (bands: axis, first and last plane) with the current sources
    // of those planes (exc_def[1])
    Exc exc_def[2];
    struct Band {
        int axis;
        unsigned lo, hi;
    };
    std::vector<Band> bands;
    // and at single cells: the currents reading a cell whose deferred source
    // changes it after the sweep, outside the bands (linear cell indices)
    unsigned long long* cell_list = nullptr;
    unsigned ncells = 0;
    // lumped RLC and dispersive lists
    std::vector<RlcDev> rlc;
    std::vector<LorDev> lor;
    std::vector<int> lor_v, lor_i;

    // probes
    std::vector<ProbeDev> probes;
    std::vector<unsigned long long> h_elin;
    std::vector<signed char> h_esign;
    std::vector<double> h_ediv;
    ProbeDev* d_probes = nullptr;
    unsigned long long* d_elin = nullptr;
    signed char* d_esign = nullptr;
    double* d_ediv = nullptr;
    bool probes_dirty = false;
    unsigned min_interval = 0;

    // field dumps
    std::vector<DumpDev> dumps;
    std::vector<unsigned> dump_npts;
    bool sampled0 = false;          // the samples due before the first step were taken

    double* energy_part = nullptr;

    // steady-state detection
    bool has_ss = false;
    SteadyDev ss{};
    double* ss_part = nullptr;

    // timestep chunks replayed as HIP graphs: captured per launch pattern
    // (the probe / dump / steady-state decisions of the chunk's steps and the
    // buffer parity of the single-sweep forms), instantiated once, replayed
    struct GraphEnt {
        std::vector<unsigned char> key;
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        size_t nalloc = 0;                // device allocations of the engine when captured
    };
    unsigned graph_chunk = 32;            // requested steps per graph (0: direct launches)
    unsigned graph_len = 0;               // steps per graph of the captured set
    std::vector<GraphEnt> graphs;
    unsigned graph_replays = 0, graph_captures = 0;

    // timing
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0, last_volt_ms = 0, last_curr_ms = 0;
    bool profile = false;
    std::vector<hipEvent_t> kev;

    template <typename T>
    int alloc(T** p, size_t n)
    {
        void* q = nullptr;
        hipError_t e = hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T));
        if (e != hipSuccess) return fail(EMS_E_NOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
        allocs.push_back(q);
        *p = (T*)q;
        return 0;
    }
    unsigned long long lin(unsigned x, unsigned y, unsigned z) const
    {
        return x + (unsigned long long)N[0] * (y + (unsigned long long)N[1] * z);
    }
};

namespace {

void parallel_rows(size_t n, const std::function<void(size_t, size_t)>& fn)
{
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (n < 64 || nt == 1) {
        fn(0, n);
        return;
    }
    std::vector<std::thread> th;
    size_t per = (n + nt - 1) / nt;
    for (unsigned t = 0; t < nt; ++t) {
        size_t a = t * per, b = std::min(n, a + per);
        if (a >= b) break;
        th.emplace_back(fn, a, b);
    }
    for (auto& t : th) t.join();
}

// N-I-J-K host array -> device SoA (x fastest) staging buffer
void nijk_to_soa(const float* src, float* dst, const unsigned* N, unsigned long long cs)
{
    const size_t rows = (size_t)3 * N[1] * N[2];
    parallel_rows(rows, [&](size_t a, size_t b) {
        for (size_t r = a; r < b; ++r) {
            const unsigned n = (unsigned)(r / ((size_t)N[1] * N[2]));
            const size_t yz = r % ((size_t)N[1] * N[2]);
            const unsigned z = (unsigned)(yz / N[1]), y = (unsigned)(yz % N[1]);
            float* d = dst + n * cs + (size_t)N[0] * (y + (size_t)N[1] * z);
            for (unsigned x = 0; x < N[0]; ++x) d[x] = src[(((size_t)n * N[0] + x) * N[1] + y) * N[2] + z];
        }
    });
}

void soa_to_nijk(const float* src, float* dst, const unsigned* N, unsigned long long cs)
{
    const size_t rows = (size_t)3 * N[1] * N[2];
    parallel_rows(rows, [&](size_t a, size_t b) {
        for (size_t r = a; r < b; ++r) {
            const unsigned n = (unsigned)(r / ((size_t)N[1] * N[2]));
            const size_t yz = r % ((size_t)N[1] * N[2]);
            const unsigned z = (unsigned)(yz / N[1]), y = (unsigned)(yz % N[1]);
            const float* s = src + n * cs + (size_t)N[0] * (y + (size_t)N[1] * z);
            for (unsigned x = 0; x < N[0]; ++x) dst[(((size_t)n * N[0] + x) * N[1] + y) * N[2] + z] = s[x];
        }
    });
}

int upload_nijk(ems_engine* h, float* dptr, const float* host)
{
    std::vector<float> stage(3 * h->cs, 0.0f);
    nijk_to_soa(host, stage.data(), h->N, h->cs);
    HIPCHK(hipMemcpy(dptr, stage.data(), stage.size() * sizeof(float), hipMemcpyHostToDevice));
    return 0;
}

template <typename T>
int upload(ems_engine* h, T** d, const T* src, size_t n)
{
    int rc = h->alloc(d, n);
    if (rc) return rc;
    if (n) HIPCHK(hipMemcpy(*d, src, n * sizeof(T), hipMemcpyHostToDevice));
    return 0;
}

unsigned choose_zdepth(const unsigned* N)
{
    // z-sweep depth: 4 planes per thread measured best on MI355X (256^3 UPML:
    // 3 and 6 within 1 %, 8: +1.4 %, 16: +5 %, 32: +14 % step time; 1024 x
    // 1024 x 768 Mur: 8 +0.4 %); shallower sweeps when the grid would
    // otherwise give fewer than 2048 workgroups for 256 CUs
    const unsigned plane = N[0] * N[1];
    const unsigned qt = (plane + kBlock - 1) / kBlock;
    unsigned d = 4;
    while (d > 1 && (unsigned long long)qt * ((N[2] + d - 1) / d) < 2048ull) d /= 2;
    return d;
}

}  // namespace

namespace {

// per-device address of `scratch`, taken once per device with that device
// current (each device holds its own copy of the array at its own address)
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

int scratch_setup(int device)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (device < 0 || device >= kMaxDevices) return fail(EMS_E_ARG, "device ordinal beyond the scratch table");
    if (g_scratch_ptr[device]) return 0;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(EMS_E_HIP, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    g_scratch_ptr[device] = static_cast<double*>(p);
    return 0;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

const char* ems_last_error(void) { return g_err.c_str(); }

int ems_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int ems_create_layout(ems_engine** out, const unsigned numLines[3], const float* vv, const float* vi,
                      const float* ii, const float* iv, int device, int layout)
{
    if (!out || !numLines || !vv || !vi || !ii || !iv) return fail(EMS_E_ARG, "null argument");
    *out = nullptr;
    if (layout < EMS_LAYOUT_AUTO || layout > EMS_LAYOUT_COMPRESSED) return fail(EMS_E_ARG, "unknown layout");
    for (int n = 0; n < 3; ++n)
        if (numLines[n] < 3) return fail(EMS_E_ARG, "at least 3 mesh lines per direction are required");
    const unsigned long long plane = (unsigned long long)numLines[0] * numLines[1];
    if (plane > 0xFFFFFFFFull || plane * numLines[2] > 0xFFFFFFFFull)
        return fail(EMS_E_ARG, "mesh too large for 32-bit in-plane indexing");
    int ndev = ems_device_count();
    if (ndev <= 0) return fail(EMS_E_NODEV, "no HIP device available");
    if (device < 0 || device >= ndev) return fail(EMS_E_ARG, "invalid device ordinal");
    HIPCHK(hipSetDevice(device));
    if (int src = scratch_setup(device)) return src;
    const unsigned long long tot = plane * numLines[2];
    const unsigned long long cs = (tot + 63) / 64 * 64;
    // per-cell compression (host): first-seen order x -> y -> z
    std::vector<uint16_t> cidx;
    std::vector<float> tab[2];
    size_t U = 0;
    if (layout != EMS_LAYOUT_PLAIN) {
        U = ems::compress_cells(numLines, vv, vi, ii, iv, cs, cidx, tab[0], tab[1], ems::host_threads());
        if (U == 0 && layout == EMS_LAYOUT_COMPRESSED)
            return fail(EMS_E_ARG, "the operator has more than 65536 unique coefficient tuples");
        if (layout == EMS_LAYOUT_AUTO && U * 16 > tot) U = 0;
    }
    ems_engine* h = new ems_engine();
    h->device = device;
    h->layout_req = layout;
    for (int n = 0; n < 3; ++n) h->N[n] = numLines[n];
    h->cs = cs;
    GridDesc& g = h->g;
    g.Nx = numLines[0];
    g.Ny = numLines[1];
    g.Nz = numLines[2];
    g.plane = (unsigned)plane;
    g.cs = h->cs;
    g.zdepth = choose_zdepth(h->N);
    if (const char* zd = std::getenv("EMS_ZDEPTH")) {
        const int v = std::atoi(zd);
        if (v >= 1 && v <= 1024) g.zdepth = (unsigned)v;
    }
    set_tiles(g, 0);
    // fields larger than the last-level cache (256 MB on MI355X) stream through it
    h->nt_fields = 2ull * 3ull * h->cs * sizeof(float) > (192ull << 20);
    if (const char* nt = std::getenv("EMS_NT_FIELDS")) h->nt_fields = std::atoi(nt) != 0;
    if (const char* qe = std::getenv("EMS_QUAD")) h->quad_req = std::atoi(qe) ? -1 : 0;
    if (const char* gc = std::getenv("EMS_GRAPH_CHUNK")) {
        const int v = std::atoi(gc);
        if (v >= 0 && v <= 1 << 20) h->graph_chunk = (unsigned)v;
    }
    int rc = 0;
    auto bail = [&](int r) {
        ems_destroy(h);
        return r;
    };
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) return bail(fail(EMS_E_HIP, "stream creation failed"));
    if ((rc = h->alloc(&h->V, 3 * h->cs)) || (rc = h->alloc(&h->I, 3 * h->cs)) || (rc = h->alloc(&h->numts_dev, 4)))
        return bail(rc);
    if (hipMemset(h->V, 0, 3 * h->cs * sizeof(float)) != hipSuccess ||
        hipMemset(h->I, 0, 3 * h->cs * sizeof(float)) != hipSuccess ||
        hipMemset(h->numts_dev, 0, 4 * sizeof(unsigned)) != hipSuccess)
        return bail(fail(EMS_E_HIP, "hipMemset failed"));
    if (U > 0) {
        h->cmp = true;
        h->n_unique = (unsigned)U;
        if ((rc = upload(h, &h->cidx, (const unsigned short*)cidx.data(), cidx.size())) ||
            (rc = upload(h, &h->ctab[0], tab[0].data(), tab[0].size())) ||
            (rc = upload(h, &h->ctab[1], tab[1].data(), tab[1].size())))
            return bail(rc);
    } else {
        if ((rc = h->alloc(&h->vv, 3 * h->cs)) || (rc = h->alloc(&h->vi, 3 * h->cs)) ||
            (rc = h->alloc(&h->ii, 3 * h->cs)) || (rc = h->alloc(&h->iv, 3 * h->cs)))
            return bail(rc);
        if ((rc = upload_nijk(h, h->vv, vv)) || (rc = upload_nijk(h, h->vi, vi)) || (rc = upload_nijk(h, h->ii, ii)) ||
            (rc = upload_nijk(h, h->iv, iv)))
            return bail(rc);
    }
    h->kV = CoefArgs{h->vv, h->vi, h->cidx, (const float4*)h->ctab[0], 2u * h->n_unique};
    h->kI = CoefArgs{h->ii, h->iv, h->cidx, (const float4*)h->ctab[1], 2u * h->n_unique};
    hipEventCreate(&h->ev0);
    hipEventCreate(&h->ev1);
    *out = h;
    return 0;
}

int ems_create(ems_engine** out, const unsigned numLines[3], const float* vv, const float* vi, const float* ii,
               const float* iv, int device)
{
    return ems_create_layout(out, numLines, vv, vi, ii, iv, device, EMS_LAYOUT_AUTO);
}

int ems_create_sse_compressed(ems_engine** out, const unsigned numLines[3], const unsigned* op_index,
                              unsigned n_unique, const float* const vv4[3], const float* const vi4[3],
                              const float* const ii4[3], const float* const iv4[3], int device, int layout)
{
    if (!out || !numLines) return fail(EMS_E_ARG, "null argument");
    *out = nullptr;
    const size_t M = (size_t)numLines[0] * numLines[1] * numLines[2];
    std::vector<float> a(3 * M), b(3 * M), c(3 * M), d(3 * M);
    const int rc = ems_expand_sse_compressed(numLines, op_index, n_unique, vv4, vi4, ii4, iv4, a.data(), b.data(),
                                             c.data(), d.data());
    if (rc) return rc;
    return ems_create_layout(out, numLines, a.data(), b.data(), c.data(), d.data(), device, layout);
}

int ems_compress_operator(const unsigned numLines[3], const float* vv, const float* vi, const float* ii,
                          const float* iv, unsigned short* idx, float* tabV, float* tabI, unsigned max_unique)
{
    if (!numLines || !vv || !vi || !ii || !iv || !idx || !tabV || !tabI) return fail(EMS_E_ARG, "null argument");
    const unsigned long long tot = (unsigned long long)numLines[0] * numLines[1] * numLines[2];
    if (tot == 0) return fail(EMS_E_ARG, "empty mesh");
    std::vector<uint16_t> ci;
    std::vector<float> tv, ti;
    const size_t U = ems::compress_cells(numLines, vv, vi, ii, iv, tot, ci, tv, ti, ems::host_threads());
    if (U == 0 || U > max_unique) return fail(EMS_E_ARG, "more unique coefficient tuples than the table holds");
    std::memcpy(idx, ci.data(), tot * sizeof(uint16_t));
    std::memcpy(tabV, tv.data(), tv.size() * sizeof(float));
    std::memcpy(tabI, ti.data(), ti.size() * sizeof(float));
    return (int)U;
}

int ems_expand_sse_compressed(const unsigned numLines[3], const unsigned* op_index, unsigned n_unique,
                              const float* const vv4[3], const float* const vi4[3], const float* const ii4[3],
                              const float* const iv4[3], float* vv, float* vi, float* ii, float* iv)
{
    if (!numLines || !op_index || !vv4 || !vi4 || !ii4 || !iv4 || !vv || !vi || !ii || !iv)
        return fail(EMS_E_ARG, "null argument");
    for (int n = 0; n < 3; ++n)
        if (!vv4[n] || !vi4[n] || !ii4[n] || !iv4[n]) return fail(EMS_E_ARG, "null compressed table");
    if (n_unique == 0) return fail(EMS_E_ARG, "empty compressed table");
    const size_t Nx = numLines[0], Ny = numLines[1], Nz = numLines[2], nV = (Nz + 3) / 4;
    for (size_t xy = 0; xy < Nx * Ny; ++xy)
        for (size_t v = 0; v < nV; ++v)
            if (op_index[xy * Nz + v] >= n_unique) return fail(EMS_E_ARG, "compressed index outside the table");
    ems::expand_sse_compressed(numLines, op_index, n_unique, vv4, vi4, ii4, iv4, vv, vi, ii, iv, ems::host_threads());
    return 0;
}

int ems_get_layout(const ems_engine* h, int* layout, unsigned* n_unique, int* pml_layout, unsigned* n_unique_pml)
{
    if (!h) return fail(EMS_E_ARG, "null handle");
    if (layout) *layout = h->cmp ? EMS_LAYOUT_COMPRESSED : EMS_LAYOUT_PLAIN;
    if (n_unique) *n_unique = h->n_unique;
    if (pml_layout)
        *pml_layout = h->pml_mode == kPmlCompressed ? EMS_LAYOUT_COMPRESSED
                      : h->pml_mode == kPmlPlain    ? EMS_LAYOUT_PLAIN
                                                    : 0;
    if (n_unique_pml) *n_unique_pml = h->n_unique_pml;
    return 0;
}

}  // extern "C"

// entries (list order) grouped per (component, cell) in order of first
// appearance, as k_excite applies them, uploaded into E (device lists only)
static int upload_exc_groups(ems_engine* h, const std::vector<ExcEnt>& ent, ems_engine::Exc& E)
{
    const unsigned K = (unsigned)ent.size();
    E.K = K;
    E.G = 0;
    if (!K) return 0;
    std::unordered_map<unsigned long long, unsigned> gid;
    std::vector<unsigned long long> glin;
    std::vector<unsigned> grp(K);
    for (unsigned k = 0; k < K; ++k) {
        const unsigned long long l = ent[k].dir * h->cs + h->lin(ent[k].x, ent[k].y, ent[k].z);
        auto it = gid.find(l);
        if (it == gid.end()) {
            it = gid.emplace(l, (unsigned)glin.size()).first;
            glin.push_back(l);
        }
        grp[k] = it->second;
    }
    const unsigned G = (unsigned)glin.size();
    std::vector<unsigned> gbeg(G + 1, 0), order(K);
    for (unsigned k = 0; k < K; ++k) ++gbeg[grp[k] + 1];
    for (unsigned g = 0; g < G; ++g) gbeg[g + 1] += gbeg[g];
    std::vector<unsigned> fill(gbeg.begin(), gbeg.end() - 1);
    for (unsigned k = 0; k < K; ++k) order[fill[grp[k]]++] = k;
    std::vector<float> ga(K);
    std::vector<unsigned> gd(K);
    for (unsigned k = 0; k < K; ++k) {
        ga[k] = ent[order[k]].amp;
        gd[k] = ent[order[k]].delay;
    }
    E.G = G;
    int rc;
    if ((rc = upload(h, &E.gbeg, gbeg.data(), G + 1)) || (rc = upload(h, &E.glin, glin.data(), G)) ||
        (rc = upload(h, &E.amp, ga.data(), K)) || (rc = upload(h, &E.delay, gd.data(), K)))
        return rc;
    return 0;
}

extern "C" {

int ems_set_excitation(ems_engine* h, const float* sig_volt, const float* sig_curr, unsigned length, int period_steps,
                       unsigned nvolt, const unsigned* vidx, const unsigned short* vdir, const float* vamp,
                       const unsigned* vdelay, unsigned ncurr, const unsigned* cidx, const unsigned short* cdir,
                       const float* camp, const unsigned* cdelay)
{
    if (!h) return fail(EMS_E_ARG, "null handle");
    if (h->inited) return fail(EMS_E_STATE, "engine already initialised");
    if (h->has_exc) return fail(EMS_E_STATE, "excitation already set");
    if (length == 0 || !sig_volt || !sig_curr) return fail(EMS_E_ARG, "empty excitation signal");
    HIPCHK(hipSetDevice(h->device));
    int rc;
    if ((rc = upload(h, &h->sig[0], sig_volt, length)) || (rc = upload(h, &h->sig[1], sig_curr, length))) return rc;
    h->sigLen = length;
    h->period = period_steps;
    const unsigned Ks[2] = {nvolt, ncurr};
    const unsigned* idx[2] = {vidx, cidx};
    const unsigned short* dir[2] = {vdir, cdir};
    const float* amp[2] = {vamp, camp};
    const unsigned* del[2] = {vdelay, cdelay};
    for (int c = 0; c < 2; ++c) {
        const unsigned K = Ks[c];
        auto& E = h->exc[c];
        E.K = K;
        if (!K) continue;
        if (!idx[c] || !dir[c] || !amp[c] || !del[c]) return fail(EMS_E_ARG, "null excitation list");
        // group the entries by (component, cell) in order of first appearance
        std::unordered_map<unsigned long long, unsigned> gid;
        std::vector<unsigned long long> glin;
        std::vector<unsigned> grp(K);
        E.host.resize(K);
        for (unsigned k = 0; k < K; ++k) {
            const unsigned x = idx[c][k], y = idx[c][K + k], z = idx[c][2 * K + k];
            if (x >= h->N[0] || y >= h->N[1] || z >= h->N[2] || dir[c][k] > 2)
                return fail(EMS_E_ARG, "excitation cell outside the mesh");
            const unsigned long long l = dir[c][k] * h->cs + h->lin(x, y, z);
            auto it = gid.find(l);
            if (it == gid.end()) {
                it = gid.emplace(l, (unsigned)glin.size()).first;
                glin.push_back(l);
            }
            grp[k] = it->second;
            E.host[k] = ExcEnt{(unsigned short)std::min(x, 65535u), (unsigned short)std::min(y, 65535u),
                               (unsigned short)std::min(z, 65535u), dir[c][k], amp[c][k], del[c][k]};
        }
        const unsigned G = (unsigned)glin.size();
        std::vector<unsigned> gbeg(G + 1, 0), order(K);
        for (unsigned k = 0; k < K; ++k) ++gbeg[grp[k] + 1];
        for (unsigned g = 0; g < G; ++g) gbeg[g + 1] += gbeg[g];
        std::vector<unsigned> fill(gbeg.begin(), gbeg.end() - 1);
        for (unsigned k = 0; k < K; ++k) order[fill[grp[k]]++] = k;   // stable: list order inside a group
        std::vector<float> ga(K);
        std::vector<unsigned> gd(K);
        for (unsigned k = 0; k < K; ++k) {
            ga[k] = amp[c][order[k]];
            gd[k] = del[c][order[k]];
        }
        E.G = G;
        E.h_gbeg = gbeg;
        E.h_glin = glin;
        E.h_amp = ga;
        E.h_delay = gd;
        if ((rc = upload(h, &E.gbeg, gbeg.data(), G + 1)) || (rc = upload(h, &E.glin, glin.data(), G)) ||
            (rc = upload(h, &E.amp, ga.data(), K)) || (rc = upload(h, &E.delay, gd.data(), K)))
            return rc;
    }
    h->has_exc = true;
    return 0;
}

int ems_add_upml(ems_engine* h, const unsigned start[3], const unsigned num[3], const float* vv, const float* vvfo,
                 const float* vvfn, const float* ii, const float* iifo, const float* iifn)
{
    if (!h || !start || !num || !vv || !vvfo || !vvfn || !ii || !iifo || !iifn) return fail(EMS_E_ARG, "null argument");
    if (h->inited) return fail(EMS_E_STATE, "engine already initialised");
    if (h->slabs.size() >= kMaxSlabs) return fail(EMS_E_ARG, "at most 6 UPML slabs");
    for (int n = 0; n < 3; ++n)
        if (num[n] == 0 || start[n] + num[n] > h->N[n]) return fail(EMS_E_ARG, "UPML slab outside the mesh");
    SlabTab t{};
    t.x0 = start[0];
    t.y0 = start[1];
    t.z0 = start[2];
    t.x1 = start[0] + num[0] - 1;
    t.y1 = start[1] + num[1] - 1;
    t.z1 = start[2] + num[2] - 1;
    t.xo = (int)(start[0] & ~3u);
    t.nx = ((start[0] + num[0] + 3u) & ~3u) - (unsigned)t.xo;
    t.ny = num[1];
    for (const auto& o : h->slabs) {
        const SlabTab& u = o.tab;
        const bool overlap = t.x0 <= u.x1 && t.x1 >= u.x0 && t.y0 <= u.y1 && t.y1 >= u.y0 && t.z0 <= u.z1 && t.z1 >= u.z0;
        if (overlap) return fail(EMS_E_ARG, "overlapping UPML slabs are not supported");
    }
    ems_engine::SlabHost sh;
    sh.tab = t;
    sh.cnt = (unsigned long long)num[0] * num[1] * num[2];
    for (int n = 0; n < 3; ++n) sh.num[n] = num[n];
    // repack [3][nx][ny][nz] (z fastest) -> [3][nz][ny][nx] (x fastest)
    auto repack = [&](const float* src, std::vector<float>& d) {
        d.resize(3 * sh.cnt);
        for (int n = 0; n < 3; ++n)
            for (unsigned x = 0; x < num[0]; ++x)
                for (unsigned y = 0; y < num[1]; ++y)
                    for (unsigned z = 0; z < num[2]; ++z)
                        d[n * sh.cnt + x + (unsigned long long)num[0] * (y + (unsigned long long)num[1] * z)] =
                            src[(((size_t)n * num[0] + x) * num[1] + y) * num[2] + z];
    };
    repack(vv, sh.v[0]);
    repack(vvfo, sh.v[1]);
    repack(vvfn, sh.v[2]);
    repack(ii, sh.i[0]);
    repack(iifo, sh.i[1]);
    repack(iifn, sh.i[2]);
    h->slabs.push_back(std::move(sh));
    return 0;
}

static int upload_upml(ems_engine* h)
{
    // All slabs' positions are numbered consecutively, each slab from a
    // multiple of 4 and padded along x to whole groups of four cells
    // (SlabTab); every array is [k][npos], padding positions hold zeros.
    // Plain: coef[9][npos] per half step.  Compressed: one index array into
    // a table of the unique 18-float tuples (V: c0, cfo, cfn; I: likewise),
    // numbered in first-seen order over the cells slab after slab, x fastest
    // (padding positions take entry 0).
    if (h->slabs.empty()) return 0;
    unsigned long long npos = 0, ncell = 0;
    std::vector<unsigned long long> first;      // first cell of each slab (compact count)
    for (auto& s : h->slabs) {
        npos = (npos + 3) & ~3ull;
        s.tab.off = npos;
        npos += (unsigned long long)s.tab.nx * s.num[1] * s.num[2];
        first.push_back(ncell);
        ncell += s.cnt;
    }
    if (npos >= 0xFFFFFFFFull) return fail(EMS_E_ARG, "too many UPML positions for 32-bit numbering");
    // position of cell j (compact, x fastest) of slab k
    auto pos_of = [&](size_t k, unsigned long long j) {
        const auto& S = h->slabs[k];
        const unsigned long long xs = j % S.num[0], r = j / S.num[0];
        const unsigned long long ys = r % S.num[1], zs = r / S.num[1];
        return S.tab.off + (unsigned long long)(S.tab.x0 - S.tab.xo) + xs + S.tab.nx * (ys + (unsigned long long)S.tab.ny * zs);
    };
    std::vector<uint32_t> pidx;
    std::vector<float> uniq;
    size_t U = 0;
    if (h->layout_req != EMS_LAYOUT_PLAIN) {
        auto get = [&](size_t i, float* t) {
            size_t s = 0;
            while (s + 1 < first.size() && i >= first[s + 1]) ++s;
            const auto& S = h->slabs[s];
            const unsigned long long j = i - first[s], cnt = S.cnt;
            for (int a = 0; a < 3; ++a)
                for (int n = 0; n < 3; ++n) {
                    t[3 * a + n] = S.v[a][n * cnt + j];
                    t[9 + 3 * a + n] = S.i[a][n * cnt + j];
                }
        };
        U = ems::dedup_tuples<18>(ncell, get, pidx, uniq, 65536, ems::host_threads());
        if (h->layout_req == EMS_LAYOUT_AUTO && U * 16 > ncell) U = 0;
    }
    const bool cmp = U > 0;
    int rc;
    unsigned* dpidx = nullptr;
    if (cmp) {
        std::vector<uint32_t> ppos(npos, 0u);
        for (size_t k = 0; k < h->slabs.size(); ++k)
            for (unsigned long long j = 0; j < h->slabs[k].cnt; ++j) ppos[pos_of(k, j)] = pidx[first[k] + j];
        if ((rc = upload(h, &dpidx, ppos.data(), ppos.size()))) return rc;
    }
    for (int side = 0; side < 2; ++side) {
        PmlArgs& P = side ? h->pmlI : h->pmlV;
        P = PmlArgs{};
        P.n = (int)h->slabs.size();
        P.npos = npos;
        for (size_t k = 0; k < h->slabs.size(); ++k) P.tab[k] = h->slabs[k].tab;
        if (cmp) {
            std::vector<float> pt(U * 12, 0.0f);
            for (size_t u = 0; u < U; ++u)
                for (int q = 0; q < 9; ++q) pt[u * 12 + q] = uniq[u * 18 + 9 * side + q];
            float* dpt;
            if ((rc = upload(h, &dpt, pt.data(), pt.size()))) return rc;
            P.ptab = (const float4*)dpt;
            P.pidx = dpidx;
        } else {
            std::vector<float> stage(9 * npos, 0.0f);
            for (size_t k = 0; k < h->slabs.size(); ++k) {
                const auto& S = h->slabs[k];
                const auto* src = side ? S.i : S.v;
                for (unsigned long long j = 0; j < S.cnt; ++j) {
                    const unsigned long long q = pos_of(k, j);
                    for (int a = 0; a < 3; ++a)
                        for (int n = 0; n < 3; ++n) stage[(3 * a + n) * npos + q] = src[a][n * S.cnt + j];
                }
            }
            float* dc;
            if ((rc = upload(h, &dc, stage.data(), stage.size()))) return rc;
            P.coef = dc;
        }
        float* fl;
        if ((rc = h->alloc(&fl, 3 * npos))) return rc;
        HIPCHK(hipMemset(fl, 0, 3 * npos * sizeof(float)));
        P.flux_in = fl;
        P.flux_out = fl;
    }
    h->pml_mode = cmp ? kPmlCompressed : kPmlPlain;
    h->n_unique_pml = cmp ? (unsigned)U : 0;
    h->slabs.clear();
    return 0;
}

int ems_add_mur(ems_engine* h, int ny, unsigned line, unsigned line_shift, const float* coeff_nyP,
                const float* coeff_nyPP, unsigned start_ts)
{
    if (!h || !coeff_nyP || !coeff_nyPP) return fail(EMS_E_ARG, "null argument");
    if (h->inited) return fail(EMS_E_STATE, "engine already initialised");
    if (ny < 0 || ny > 2 || line >= h->N[ny] || line_shift >= h->N[ny]) return fail(EMS_E_ARG, "bad Mur face");
    HIPCHK(hipSetDevice(h->device));
    const int nyP = (ny + 1) % 3, nyPP = (ny + 2) % 3;
    MurDev m{};
    m.na = h->N[nyP];
    m.nb = h->N[nyPP];
    m.offP = nyP * h->cs;
    m.offPP = nyPP * h->cs;
    m.start_ts = start_ts;
    const size_t cnt = (size_t)m.na * m.nb;
    // device order of the face points: the lower tangential axis fastest (x
    // for y and z faces, y for x faces), so that neighbouring threads touch
    // neighbouring field elements; the caller's arrays are [nyP][nyPP]
    const bool tr = nyP < nyPP;
    std::vector<unsigned> face(cnt), shift(cnt);
    std::vector<float> cP(cnt), cPP(cnt);
    for (unsigned i = 0; i < m.na; ++i)
        for (unsigned j = 0; j < m.nb; ++j) {
            unsigned p[3], s[3];
            p[ny] = line;
            s[ny] = line_shift;
            p[nyP] = s[nyP] = i;
            p[nyPP] = s[nyPP] = j;
            const size_t src = (size_t)i * m.nb + j, k = tr ? (size_t)j * m.na + i : src;
            face[k] = (unsigned)h->lin(p[0], p[1], p[2]);
            shift[k] = (unsigned)h->lin(s[0], s[1], s[2]);
            cP[k] = coeff_nyP[src];
            cPP[k] = coeff_nyPP[src];
        }
    int rc;
    unsigned *df, *dsh;
    float *dcp, *dcpp;
    if ((rc = upload(h, &df, face.data(), cnt)) || (rc = upload(h, &dsh, shift.data(), cnt)) ||
        (rc = upload(h, &dcp, cP.data(), cnt)) || (rc = upload(h, &dcpp, cPP.data(), cnt)) ||
        (rc = h->alloc(&m.mP, cnt)) || (rc = h->alloc(&m.mPP, cnt)))
        return rc;
    HIPCHK(hipMemset(m.mP, 0, cnt * sizeof(float)));
    HIPCHK(hipMemset(m.mPP, 0, cnt * sizeof(float)));
    m.face = df;
    m.shift = dsh;
    m.cP = dcp;
    m.cPP = dcpp;
    h->mur.push_back(m);
    h->mur_geo.push_back(ems_engine::MurGeo{ny, line, line_shift});
    return 0;
}

int ems_add_lumped_rlc(ems_engine* h, unsigned count, const unsigned* pos, const int* dir, const float* const coeffs[9])
{
    if (!h || !pos || !dir || !coeffs) return fail(EMS_E_ARG, "null argument");
    if (h->inited) return fail(EMS_E_STATE, "engine already initialised");
    if (!count) return 0;
    HIPCHK(hipSetDevice(h->device));
    RlcDev r{};
    r.count = count;
    std::vector<unsigned long long> lin(count);
    for (unsigned k = 0; k < count; ++k) {
        const unsigned x = pos[k], y = pos[count + k], z = pos[2 * count + k];
        if (x >= h->N[0] || y >= h->N[1] || z >= h->N[2] || dir[k] < 0 || dir[k] > 2)
            return fail(EMS_E_ARG, "lumped element outside the mesh");
        lin[k] = dir[k] * h->cs + h->lin(x, y, z);
    }
    int rc;
    unsigned long long* dl;
    if ((rc = upload(h, &dl, lin.data(), count))) return rc;
    r.lin = dl;
    for (int c = 0; c < 9; ++c) {
        float* d;
        if (!coeffs[c]) return fail(EMS_E_ARG, "null RLC coefficient array");
        if ((rc = upload(h, &d, coeffs[c], count))) return rc;
        r.c[c] = d;
    }
    if ((rc = h->alloc(&r.vd, 3 * count)) || (rc = h->alloc(&r.jn, 3 * count)) || (rc = h->alloc(&r.il, count)))
        return rc;
    HIPCHK(hipMemset(r.vd, 0, 3 * count * sizeof(float)));
    HIPCHK(hipMemset(r.jn, 0, 3 * count * sizeof(float)));
    HIPCHK(hipMemset(r.il, 0, count * sizeof(float)));
    h->rlc.push_back(r);
    return 0;
}

int ems_add_lorentz(ems_engine* h, unsigned count, const unsigned* pos, const float* v_int, const float* v_ext,
                    const float* v_lor, const float* i_int, const float* i_ext, const float* i_lor, int volt_on,
                    int curr_on)
{
    if (!h || !pos) return fail(EMS_E_ARG, "null argument");
    if (h->inited) return fail(EMS_E_STATE, "engine already initialised");
    if (!count) return 0;
    HIPCHK(hipSetDevice(h->device));
    std::vector<unsigned> lin(count);
    for (unsigned k = 0; k < count; ++k) {
        const unsigned x = pos[k], y = pos[count + k], z = pos[2 * count + k];
        if (x >= h->N[0] || y >= h->N[1] || z >= h->N[2]) return fail(EMS_E_ARG, "dispersive cell outside the mesh");
        lin[k] = (unsigned)h->lin(x, y, z);
    }
    int rc;
    unsigned* dl;
    if ((rc = upload(h, &dl, lin.data(), count))) return rc;
    for (int side = 0; side < 2; ++side) {
        const bool on = side ? curr_on : volt_on;
        LorDev d{};
        d.count = count;
        d.lin = dl;
        const float* ci = side ? i_int : v_int;
        const float* ce = side ? i_ext : v_ext;
        const float* cl = side ? i_lor : v_lor;
        if (on) {
            if (!ci || !ce) return fail(EMS_E_ARG, "null ADE coefficient array");
            float *a, *b, *c = nullptr;
            if ((rc = upload(h, &a, ci, 3 * count)) || (rc = upload(h, &b, ce, 3 * count))) return rc;
            if (cl && (rc = upload(h, &c, cl, 3 * count))) return rc;
            d.cint = a;
            d.cext = b;
            d.clor = c;
            if ((rc = h->alloc(&d.ade, 3 * count)) || (rc = h->alloc(&d.lor, 3 * count))) return rc;
            HIPCHK(hipMemset(d.ade, 0, 3 * count * sizeof(float)));
            HIPCHK(hipMemset(d.lor, 0, 3 * count * sizeof(float)));
        }
        h->lor.push_back(d);
        (side ? h->lor_i : h->lor_v).push_back(on ? 1 : 0);
    }
    return 0;
}

// z-sweep depth of the fused step: 16 planes per workgroup, shallower when
// the grid would otherwise give fewer than 1024 workgroups
static unsigned choose_fused_zdepth(unsigned ntiles, unsigned Nz)
{
    unsigned d = 16;
    while (d > 4 && (unsigned long long)ntiles * ((Nz + d - 1) / d) < 1024ull) d /= 2;
    if (const char* zd = std::getenv("EMS_FZDEPTH")) {
        const int v = std::atoi(zd);
        if (v >= 1 && v <= 1024) d = (unsigned)v;
    }
    return d;
}

// Decide and prepare the fused single-sweep step: every active extension
// must have a fused form (UPML, excitation); Mur, Lorentz and lumped RLC keep
// the two-half-step path.  Form 1 (k_step) stages V^{n+1} planes in LDS per
// 64 x 16 tile; form 2 (k_step2) sweeps single columns and recomputes the
// neighbour components it needs.
static int setup_fused(ems_engine* h)
{
    h->fused = 0;
    // automatic mode: the tiled quad sweep (form 3, k_step4t) on meshes
    // without UPML whose rows hold at least 64 quads (1024 x 1024 x 768 Mur
    // layers 12.08 -> 11.65 ms per step, 256^3 PEC 0.218 -> 0.202 ms), where
    // every extension has a form for it; two half steps elsewhere (256^3
    // UPML, compressed: 0.50 ms per step in the LDS sweep, 0.38 ms in two
    // half steps)
    int form = h->fused_req;
    const bool autom = form == -1;
    if (autom) {
        bool sweep = !h->pmlV.n && h->quad_req != 0 && h->N[0] % 4 == 0 && h->N[0] / 4 >= 64;
        if (const char* as = std::getenv("EMS_AUTO_SWEEP")) sweep = sweep && std::atoi(as) != 0;
        if (sweep) {
            // the sweep keeps a second copy of both fields: only where it fits
            // with 1 GiB to spare
            size_t freeb = 0, totalb = 0;
            const size_t need = 6 * h->cs * sizeof(float) + (1ull << 30);
            if (hipMemGetInfo(&freeb, &totalb) != hipSuccess || freeb < need) sweep = false;
        }
        if (!sweep) return 0;
        form = 3;
    }
    if (form <= 0) return 0;
    std::string why;
    const unsigned Nx = h->N[0], Ny = h->N[1], Nz = h->N[2];
    if (!h->mur.empty() && form != 3) why = "Mur ABC has no fused form but the quad sweep's";
    else if (!h->lor.empty()) why = "Lorentz / Drude materials have no fused form";
    else if (!h->rlc.empty()) why = "lumped RLC elements have no fused form";
    else if (h->has_ss) why = "steady-state detection has no fused form";
    else if (Nx >= 65535 || Ny >= 65535 || Nz >= 65535) why = "mesh too large for the fused step";
    for (const auto& e : h->exc[1].host)
        if (why.empty() && (e.x + 1u >= Nx || e.y + 1u >= Ny || e.z + 1u >= Nz))
            why = "a current source lies outside the updated range";
    // source lists in list order: per tile (form 1: voltage sources of the
    // tile and its halo, column x0+kTX and row y0+kTY; current sources of the
    // tile) or per (x, y) column (form 2)
    const unsigned ntx = (Nx + kTX - 1) / kTX, nty = (Ny + kTY - 1) / kTY;
    if (why.empty() && form == 3) {
        if (h->pmlV.n) why = "the quad sweep has no UPML form";
        else if (Nx % 4) why = "the quad sweep needs a multiple of 4 x lines";
    }
    // quad sweep with Mur faces: voltage sources on a face or neighbour line
    // wait for the Mur hooks (exc_def[0]); the current update of the planes
    // whose voltages the hooks and those sources change is redone (bands)
    std::vector<ExcEnt> deferred[2];
    auto on_face = [&](const ExcEnt& e) {
        const unsigned p[3] = {e.x, e.y, e.z};
        for (const auto& f : h->mur_geo)
            if (p[f.ny] == f.line || p[f.ny] == f.shift) return true;
        return false;
    };
    h->bands.clear();
    h->ncells = 0;
    h->mur_fold[0] = h->mur_fold[1] = -1;
    std::vector<unsigned long long> cells;
    if (why.empty() && form == 3) {
        // k_step4t on rows of at least 64 quads (EMS_Q4TILE forces it on or off)
        h->q4tile = Nx / 4 >= 64;
        if (const char* te = std::getenv("EMS_Q4TILE")) h->q4tile = std::atoi(te) != 0;
        // it runs the pre and post hooks of the Mur faces on x = 0 and
        // x = Nx-1 and updates the currents next to them itself
        bool fold = h->q4tile && Nx >= 8;
        if (const char* mf = std::getenv("EMS_MUR_FOLD")) fold = fold && std::atoi(mf) != 0;
        for (size_t k = 0; k < h->mur_geo.size() && fold; ++k) {
            const auto& f = h->mur_geo[k];
            if (f.ny != 0) continue;
            if (f.line == 0 && f.shift == 1 && h->mur_fold[0] < 0) h->mur_fold[0] = (int)k;
            else if (f.line == Nx - 1 && f.shift == Nx - 2 && h->mur_fold[1] < 0) h->mur_fold[1] = (int)k;
        }
        // the Mur apply sets the voltages of the face line only; the current
        // at plane k reads the voltages of planes k and k+1 along the normal
        for (size_t k = 0; k < h->mur_geo.size(); ++k) {
            const auto& f = h->mur_geo[k];
            if ((int)k == h->mur_fold[0] || (int)k == h->mur_fold[1]) continue;
            const unsigned lo = f.line ? f.line - 1 : 0, hi = std::min(f.line, h->N[f.ny] - 2);
            if (lo <= hi) h->bands.push_back(ems_engine::Band{f.ny, lo, hi});
        }
        auto in_band = [&](const unsigned p[3]) {
            for (const auto& b : h->bands)
                if (p[b.axis] >= b.lo && p[b.axis] <= b.hi) return true;
            return false;
        };
        // a deferred voltage source changes its cell after the sweep: the
        // currents of the cell and of its -x, -y, -z neighbours read it
        for (const auto& e : h->exc[0].host) {
            if (!on_face(e)) continue;
            for (int d = 0; d < 4; ++d) {
                unsigned p[3] = {e.x, e.y, e.z};
                if (d < 3) {
                    if (!p[d]) continue;
                    --p[d];
                }
                if (p[0] + 1 >= Nx || p[1] + 1 >= Ny || p[2] + 1 >= Nz || in_band(p)) continue;
                cells.push_back(h->lin(p[0], p[1], p[2]));
            }
        }
        std::sort(cells.begin(), cells.end());
        cells.erase(std::unique(cells.begin(), cells.end()), cells.end());
        for (const auto& e : h->exc[1].host) {
            const unsigned p[3] = {e.x, e.y, e.z};
            if (in_band(p) || std::binary_search(cells.begin(), cells.end(), h->lin(e.x, e.y, e.z)))
                deferred[1].push_back(e);
        }
    }
    const unsigned nlists = form == 1 ? ntx * nty : h->g.plane;
    std::vector<std::vector<ExcEnt>> lists[2];
    size_t most = 0;
    for (int c = 0; c < 2 && why.empty(); ++c) {
        lists[c].assign(nlists, {});
        for (const auto& e : h->exc[c].host) {
            if (form == 3 && c == 0 && !h->mur_geo.empty() && on_face(e)) {
                deferred[0].push_back(e);
                continue;
            }
            if (form >= 2) {
                lists[c][e.x + Nx * (unsigned)e.y].push_back(e);
                continue;
            }
            const unsigned txs[2] = {e.x / kTX, (c == 0 && e.x % kTX == 0 && e.x > 0) ? e.x / kTX - 1 : ~0u};
            const unsigned tys[2] = {e.y / kTY, (c == 0 && e.y % kTY == 0 && e.y > 0) ? e.y / kTY - 1 : ~0u};
            for (unsigned tx : txs)
                for (unsigned ty : tys)
                    if (tx < ntx && ty < nty) lists[c][tx + ntx * ty].push_back(e);
        }
        for (const auto& l : lists[c]) most = std::max(most, l.size());
    }
    if (why.empty() && most > 64) why = form == 1 ? "more than 64 sources in one tile" : "more than 64 sources in one column";
    if (!why.empty()) {
        h->bands.clear();
        return autom ? 0 : fail(EMS_E_ARG, "fused step unavailable: " + why);
    }
    int rc;
    for (int c = 0; c < 2; ++c)
        if ((rc = upload_exc_groups(h, deferred[c], h->exc_def[c]))) return rc;
    if (!cells.empty()) {
        if ((rc = upload(h, &h->cell_list, cells.data(), cells.size()))) return rc;
        h->ncells = (unsigned)cells.size();
    }
    for (int side = 0; side < 2; ++side) {
        h->nrim[side] = 0;
        if (h->mur_fold[side] < 0) continue;
        // points z * Ny + y of an x face with y or z within 2 lines of the edge
        std::vector<unsigned> rim;
        for (unsigned z = 0; z < Nz; ++z)
            for (unsigned y = 0; y < Ny; ++y)
                if (y < 2 || y + 2 >= Ny || z < 2 || z + 2 >= Nz) rim.push_back(z * Ny + y);
        if ((rc = upload(h, &h->mur_rim[side], rim.data(), rim.size()))) return rc;
        h->nrim[side] = (unsigned)rim.size();
    }
    for (int c = 0; c < 2; ++c) {
        std::vector<unsigned> beg(nlists + 1, 0);
        std::vector<ExcEnt> ent;
        for (unsigned t = 0; t < nlists; ++t) {
            ent.insert(ent.end(), lists[c][t].begin(), lists[c][t].end());
            beg[t + 1] = (unsigned)ent.size();
            // form 3: the entries of each group of four columns (one quad)
            // ordered by plane, list order kept within a plane
            if (form == 3 && t % 4 == 3)
                std::stable_sort(ent.begin() + beg[t - 3], ent.end(),
                                 [](const ExcEnt& l, const ExcEnt& r) { return l.z < r.z; });
        }
        ExcArgs& E = c ? h->eI : h->eV;
        unsigned* db;
        ExcEnt* de;
        if ((rc = upload(h, &db, beg.data(), beg.size())) || (rc = upload(h, &de, ent.data(), ent.size()))) return rc;
        E.beg = db;
        E.ent = de;
        E.sig = h->sig[c];
        E.length = h->sigLen;
        E.period = h->period;
    }
    // second buffers for state n+1; cells a step does not write (the last
    // current line of each direction) must hold the same values in both
    if ((rc = h->alloc(&h->V_alt, 3 * h->cs)) || (rc = h->alloc(&h->I_alt, 3 * h->cs))) return rc;
    HIPCHK(hipMemcpy(h->V_alt, h->V, 3 * h->cs * sizeof(float), hipMemcpyDeviceToDevice));
    HIPCHK(hipMemcpy(h->I_alt, h->I, 3 * h->cs * sizeof(float), hipMemcpyDeviceToDevice));
    for (int side = 0; side < 2; ++side) {
        PmlArgs& P = side ? h->pmlI : h->pmlV;
        if (!P.n) continue;
        if ((rc = h->alloc(&h->flux_alt[side], 3 * P.npos))) return rc;
        HIPCHK(hipMemcpy(h->flux_alt[side], P.flux_in, 3 * P.npos * sizeof(float), hipMemcpyDeviceToDevice));
    }
    HIPCHK(hipDeviceSynchronize());
    h->V_base = h->V;
    if (form == 1) {
        h->ntx = ntx;
        h->nty = nty;
        h->fzdepth = choose_fused_zdepth(ntx * nty, Nz);
        h->fnzc = (Nz + h->fzdepth - 1) / h->fzdepth;
    } else if (form == 3) {
        // tiles of kBlock quads over the plane (k_step4), or of 64 quads x 4
        // rows (k_step4t, rows of at least 64 quads); 8 planes per chunk (128
        // for the tiles: 10.72, 10.61, 10.51 ms per sweep on the Mur layers
        // at 16, 32, 64, and 10.47 against 10.53 at 128), shallower when the
        // grid would give fewer than 2048 workgroups
        const unsigned NQ = Nx / 4;
        h->q4ntq = (NQ + 63) / 64;
        h->ntx = h->q4tile ? h->q4ntq * ((Ny + kQTileRows - 1) / kQTileRows) : (h->g.plane / 4 + kBlock - 1) / kBlock;
        unsigned d = h->q4tile ? 128 : 8;
        while (d > 2 && (unsigned long long)h->ntx * ((Nz + d - 1) / d) < 2048ull) d /= 2;
        if (const char* zd = std::getenv("EMS_Q4ZDEPTH")) {
            const int v = std::atoi(zd);
            if (v >= 1 && v <= 1024) d = (unsigned)v;
        }
        h->fzdepth = d;
        h->fnzc = (Nz + d - 1) / d;
    } else {
        // the two-half-step tiles (kBlock consecutive columns) and z depth
        h->fzdepth = h->g.zdepth;
        h->fnzc = h->g.nzc;
    }
    h->fused = form;
    return 0;
}

int ems_set_fused(ems_engine* h, int mode)
{
    if (!h) return fail(EMS_E_ARG, "null handle");
    if (h->inited) return fail(EMS_E_STATE, "engine already initialised");
    if (mode < -1 || mode > 3)
        return fail(EMS_E_ARG, "step form is -1 (auto), 0 (two half steps), 1 (LDS sweep) or 2 (column sweep)");
    h->fused_req = mode;
    return 0;
}

int ems_get_fused(const ems_engine* h, int* fused, unsigned* zdepth)
{
    if (!h) return fail(EMS_E_ARG, "null handle");
    if (fused) *fused = h->fused;
    if (zdepth) *zdepth = h->fused ? h->fzdepth : h->g.zdepth;
    return 0;
}

int ems_set_steady_state(ems_engine* h, unsigned period_steps, unsigned nprobe, const int* dir, const unsigned* pos,
                         unsigned history_capacity)
{
    if (!h || !dir || !pos) return fail(EMS_E_ARG, "null argument");
    if (h->inited) return fail(EMS_E_STATE, "engine already initialised");
    if (h->has_ss) return fail(EMS_E_STATE, "steady-state detection already set");
    if (period_steps == 0 || nprobe == 0) return fail(EMS_E_ARG, "empty period or probe list");
    std::vector<unsigned long long> lin(nprobe);
    for (unsigned k = 0; k < nprobe; ++k) {
        const unsigned x = pos[k], y = pos[nprobe + k], z = pos[2 * nprobe + k];
        if (dir[k] < 0 || dir[k] > 2 || x >= h->N[0] || y >= h->N[1] || z >= h->N[2])
            return fail(EMS_E_ARG, "steady-state probe outside the mesh");
        lin[k] = (unsigned long long)dir[k] * h->cs + h->lin(x, y, z);
    }
    HIPCHK(hipSetDevice(h->device));
    SteadyDev S{};
    S.n = nprobe;
    S.p = period_steps;
    S.cap = std::max(1u, history_capacity);
    unsigned long long* dl;
    int rc;
    if ((rc = upload(h, &dl, lin.data(), nprobe)) || (rc = h->alloc(&S.rec, (size_t)nprobe * 2 * period_steps)) ||
        (rc = h->alloc(&S.state, 3)) || (rc = h->alloc(&S.hist_ts, S.cap)) || (rc = h->alloc(&S.hist_diff, S.cap)) ||
        (rc = h->alloc(&h->ss_part, 2 * 1024)))
        return rc;
    S.lin = dl;
    HIPCHK(hipMemset(S.rec, 0, (size_t)nprobe * 2 * period_steps * sizeof(float)));
    const double st[3] = {1.0, 0.0, 0.0};
    HIPCHK(hipMemcpy(S.state, st, sizeof(st), hipMemcpyHostToDevice));
    h->ss = S;
    h->has_ss = true;
    return 0;
}

int ems_steady_state(ems_engine* h, double* last_max_diff, unsigned* evaluations, unsigned* ts, double* diff,
                     unsigned cap)
{
    if (!h) return fail(EMS_E_ARG, "null handle");
    if (!h->has_ss) return fail(EMS_E_STATE, "no steady-state detection");
    HIPCHK(hipSetDevice(h->device));
    HIPCHK(hipStreamSynchronize(h->stream));
    double st[3];
    HIPCHK(hipMemcpy(st, h->ss.state, sizeof(st), hipMemcpyDeviceToHost));
    if (last_max_diff) *last_max_diff = st[0];
    const unsigned n = (unsigned)st[2];
    if (evaluations) *evaluations = n;
    const unsigned m = std::min(std::min(n, h->ss.cap), cap);
    if (ts && m) HIPCHK(hipMemcpy(ts, h->ss.hist_ts, m * sizeof(unsigned), hipMemcpyDeviceToHost));
    if (diff && m) HIPCHK(hipMemcpy(diff, h->ss.hist_diff, m * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

// Quad half steps (k_volt4 / k_curr4): two-half-step engines whose rows hold
// a whole number of four-cell groups.  Automatic mode takes them on meshes
// without UPML (1024 x 1024 x 768 Mur layers: 6.37 -> 5.79 ms per half
// step).  On a UPML mesh, quads run as two launches: the box without UPML
// cells without UPML code, the six boxes around it with it.  Automatic mode
// keeps one cell per lane there: on the 256^3 UPML bricks the UPML-free box
// (82 % of the cells) takes 119 us, but the quads' UPML path over the
// remaining 18 % takes 75 us, against 161 us for the whole mesh at one cell
// per lane.  z depth: 4 planes per workgroup, shallower when the grid would
// give fewer than 2048 workgroups.
static int setup_quad(ems_engine* h)
{
    h->quad = false;
    if (h->quad_req == 0) return 0;
    std::string why;
    if (h->fused) why = "the single-sweep step forms have no quad form";
    else if (h->N[0] % 4) why = "the number of x lines is not a multiple of 4";
    if (!why.empty()) return h->quad_req == 1 ? fail(EMS_E_ARG, "quad half steps unavailable: " + why) : 0;
    GridDesc& q = h->gq;
    q = h->g;
    q.tx = 0;
    q.ntx = 0;
    q.qtiles = (q.plane / 4 + kBlock - 1) / kBlock;
    unsigned d = 4;
    while (d > 1 && (unsigned long long)q.qtiles * ((q.Nz + d - 1) / d) < 2048ull) d /= 2;
    if (const char* zd = std::getenv("EMS_QZDEPTH")) {
        const int v = std::atoi(zd);
        if (v >= 1 && v <= 1024) d = (unsigned)v;
    }
    q.zdepth = d;
    q.nzc = (q.Nz + d - 1) / d;
    const unsigned NQ = h->N[0] / 4;
    // boxes in quads x rows x planes; workgroups counted box after box
    QBox hb[3][kMaxQBoxes] = {};          // all, in, out
    unsigned nb[3] = {0, 0, 0}, blk[3] = {0, 0, 0};
    // the boxes around the UPML-free box hold few cells and the longer UPML
    // update: one plane per workgroup, for enough waves to hide its latency
    unsigned dout = 1;
    if (const char* zo = std::getenv("EMS_QZOUT")) {
        const int v = std::atoi(zo);
        if (v >= 1 && v <= 1024) dout = (unsigned)v;
    }
    auto add = [&](int set, unsigned qa, unsigned nq, unsigned ya, unsigned ny, unsigned za, unsigned zb) {
        if (!nq || !ny || za >= zb) return;
        QBox& b = hb[set][nb[set]++];
        b.qa = qa;
        b.nq = nq;
        b.ya = ya;
        b.ny = ny;
        b.za = za;
        b.zb = zb;
        b.zd = set == 2 ? dout : d;
        b.ntiles = (nq * ny + kBlock - 1) / kBlock;
        b.first = blk[set];
        blk[set] += b.ntiles * ((zb - za + b.zd - 1) / b.zd);
    };
    add(0, 0, NQ, 0, h->N[1], 0, h->N[2]);
    // UPML mesh: the largest box of whole quads holding no UPML cell, found
    // along the mesh's centre lines, and the six boxes around it
    h->qsplit = false;
    if (h->pmlV.n) {
        const PmlArgs& P = h->pmlV;
        const unsigned c[3] = {h->N[0] / 2, h->N[1] / 2, h->N[2] / 2};
        unsigned lo[3], hi[3];                  // interior [lo, hi) per axis
        for (int a = 0; a < 3; ++a) {
            auto in_slab = [&](unsigned v) {
                unsigned p[3] = {c[0], c[1], c[2]};
                p[a] = v;
                for (int s = 0; s < P.n; ++s) {
                    const SlabTab& S = P.tab[s];
                    if ((int)p[0] >= S.x0 && (int)p[0] <= S.x1 && (int)p[1] >= S.y0 && (int)p[1] <= S.y1 &&
                        (int)p[2] >= S.z0 && (int)p[2] <= S.z1)
                        return true;
                }
                return false;
            };
            lo[a] = 0;
            while (lo[a] < h->N[a] && in_slab(lo[a])) ++lo[a];
            hi[a] = h->N[a];
            while (hi[a] > lo[a] && in_slab(hi[a] - 1)) --hi[a];
        }
        const unsigned qa = (lo[0] + 3) / 4, qb = hi[0] / 4;
        bool ok = qa < qb && lo[1] < hi[1] && lo[2] < hi[2];
        // no slab may reach into the box
        for (int s = 0; s < P.n && ok; ++s) {
            const SlabTab& S = P.tab[s];
            const bool meet = S.x0 < (int)(4 * qb) && S.x1 >= (int)(4 * qa) && S.y0 < (int)hi[1] &&
                              S.y1 >= (int)lo[1] && S.z0 < (int)hi[2] && S.z1 >= (int)lo[2];
            if (meet) ok = false;
        }
        if (ok) {
            add(1, qa, qb - qa, lo[1], hi[1] - lo[1], lo[2], hi[2]);
            const unsigned Ny = h->N[1], Nz = h->N[2];
            add(2, 0, NQ, 0, Ny, 0, lo[2]);                        // planes below
            add(2, 0, NQ, 0, Ny, hi[2], Nz);                       // planes above
            add(2, 0, NQ, 0, lo[1], lo[2], hi[2]);                 // rows in front
            add(2, 0, NQ, hi[1], Ny - hi[1], lo[2], hi[2]);        // rows behind
            add(2, 0, qa, lo[1], hi[1] - lo[1], lo[2], hi[2]);     // quads left
            add(2, qb, NQ - qb, lo[1], hi[1] - lo[1], lo[2], hi[2]);  // quads right
            h->qsplit = true;
        }
    }
    if (!h->qbox_dev) {
        int rc;
        if ((rc = h->alloc(&h->qbox_dev, 3 * kMaxQBoxes))) return rc;
    }
    HIPCHK(hipMemcpy(h->qbox_dev, hb, sizeof(hb), hipMemcpyHostToDevice));
    h->qbox_all = QBoxes{h->qbox_dev, nb[0]};
    h->qbox_in = QBoxes{h->qbox_dev + kMaxQBoxes, nb[1]};
    h->qbox_out = QBoxes{h->qbox_dev + 2 * kMaxQBoxes, nb[2]};
    h->qblk_all = blk[0];
    h->qblk_in = blk[1];
    h->qblk_out = blk[2];
    if (const char* sp = std::getenv("EMS_QSPLIT")) h->qsplit = h->qsplit && std::atoi(sp) != 0;
    // automatic mode: quads on meshes without UPML
    if (h->quad_req == -1 && h->pmlV.n) return 0;
    h->quad = true;
    return 0;
}

// Voltage sources added by the half-step voltage kernel itself (k_volt4 per
// quad, k_volt per column) when no hook runs between the voltage update and
// the excitation (Mur post and apply, lumped RLC, Lorentz, the steady-state
// record): one launch less per step.  Lists per quad or column, ordered by
// plane, list order kept within a plane.
static int setup_vsrc(ems_engine* h)
{
    h->qsrc = ExcArgs{};
    bool on = !h->fused && h->exc[0].G && h->exc[0].host.size() == h->exc[0].K && h->mur.empty() &&
              h->lor.empty() && h->rlc.empty() && !h->has_ss && h->N[0] < 65535 && h->N[1] < 65535 &&
              h->N[2] < 65535;
    if (const char* qs = std::getenv("EMS_QSRC")) on = on && std::atoi(qs) != 0;
    if (!on) return 0;
    const unsigned w = h->quad ? 4u : 1u;            // cells per list along x
    const unsigned nx = h->N[0] / w;
    const unsigned nl = (unsigned)(h->g.plane / w);
    std::vector<std::vector<ExcEnt>> lists(nl);
    for (const auto& e : h->exc[0].host) lists[e.x / w + nx * (unsigned)e.y].push_back(e);
    std::vector<unsigned> beg(nl + 1, 0);
    std::vector<ExcEnt> ent;
    for (unsigned i = 0; i < nl; ++i) {
        std::stable_sort(lists[i].begin(), lists[i].end(), [](const ExcEnt& l, const ExcEnt& r) { return l.z < r.z; });
        ent.insert(ent.end(), lists[i].begin(), lists[i].end());
        beg[i + 1] = (unsigned)ent.size();
    }
    unsigned* db;
    ExcEnt* de;
    unsigned* dm = nullptr;
    int rc;
    if ((rc = upload(h, &db, beg.data(), beg.size())) || (rc = upload(h, &de, ent.data(), ent.size()))) return rc;
    if (w == 1 && h->g.tx == 0) {
        // k_volt: tiles of kBlock consecutive columns (the single engine's tile
        // shape); only tiles holding a source read the lists
        const unsigned nt = (unsigned)((h->g.plane + kBlock - 1) / kBlock);
        std::vector<unsigned> mask((nt + 31) / 32, 0u);
        for (const auto& e : h->exc[0].host) {
            const unsigned t = (unsigned)(((unsigned long long)e.x + (unsigned long long)h->N[0] * e.y) / kBlock);
            mask[t >> 5] |= 1u << (t & 31);
        }
        if ((rc = upload(h, &dm, mask.data(), mask.size()))) return rc;
    }
    h->qsrc = ExcArgs{db, de, h->sig[0], h->sigLen, h->period, dm};
    return 0;
}

int ems_set_quad(ems_engine* h, int mode)
{
    if (!h) return fail(EMS_E_ARG, "null handle");
    if (h->inited) return fail(EMS_E_STATE, "engine already initialised");
    if (mode < -1 || mode > 1) return fail(EMS_E_ARG, "quad mode is -1 (auto), 0 (off) or 1 (on)");
    h->quad_req = mode;
    return 0;
}

int ems_get_quad(const ems_engine* h, int* on, unsigned* zdepth)
{
    if (!h) return fail(EMS_E_ARG, "null handle");
    if (on) *on = h->quad ? 1 : 0;
    if (zdepth) *zdepth = h->quad ? h->gq.zdepth : 0;
    return 0;
}

int ems_init(ems_engine* h)
{
    if (!h) return fail(EMS_E_ARG, "null handle");
    if (h->inited) return 0;
    HIPCHK(hipSetDevice(h->device));
    int rc = upload_upml(h);
    if (rc) return rc;
    if ((rc = setup_fused(h))) return rc;
    if ((rc = setup_quad(h))) return rc;
    if ((rc = setup_vsrc(h))) return rc;
    // small compressed tables (both half steps' within 32 KiB) go to LDS in the quad kernels
    h->lds_tab = h->cmp && (size_t)4 * h->n_unique * sizeof(float4) <= kLdsTabBytes;
    if (const char* lt = std::getenv("EMS_LDS_TAB")) h->lds_tab = h->lds_tab && std::atoi(lt) != 0;
    h->inited = true;
    return 0;
}

static void drop_graphs(ems_engine* h);

int ems_add_probe(ems_engine* h, const ems_probe_desc* d, int* id)
{
    if (!h || !d) return fail(EMS_E_ARG, "null argument");
    if (d->type < 0 || d->type > 3) return fail(EMS_E_ARG, "unknown probe type");
    if (d->nelem == 0 || !d->comp || !d->pos) return fail(EMS_E_ARG, "probe without terms");
    if ((d->type == EMS_PROBE_EFIELD || d->type == EMS_PROBE_HFIELD) && (d->nelem != 3 || !d->divisor))
        return fail(EMS_E_ARG, "field probes take three terms with divisors");
    if ((d->type == EMS_PROBE_VOLTAGE || d->type == EMS_PROBE_CURRENT) && !d->sign)
        return fail(EMS_E_ARG, "integral probes need term signs");
    ProbeDev p{};
    p.type = d->type;
    p.nelem = d->nelem;
    p.off = (unsigned)h->h_elin.size();
    p.interval = d->interval;
    p.fd_interval = d->fd_interval;
    p.start_ts = d->start_ts;
    p.stop_ts = d->stop_ts;
    p.capacity = d->capacity;
    for (unsigned e = 0; e < d->nelem; ++e) {
        const unsigned x = d->pos[e], y = d->pos[d->nelem + e], z = d->pos[2 * d->nelem + e];
        if (x >= h->N[0] || y >= h->N[1] || z >= h->N[2] || d->comp[e] > 2)
            return fail(EMS_E_ARG, "probe term outside the mesh");
        h->h_elin.push_back(d->comp[e] * h->cs + h->lin(x, y, z));
        h->h_esign.push_back(d->sign ? (signed char)(d->sign[e] >= 0 ? 1 : -1) : 1);
        h->h_ediv.push_back(d->divisor ? d->divisor[e] : 0.0);
    }
    HIPCHK(hipSetDevice(h->device));
    int rc;
    if ((rc = h->alloc(&p.count, 1)) || (rc = h->alloc(&p.ts_out, std::max(1u, p.capacity))) ||
        (rc = h->alloc(&p.val_out, 3 * (size_t)std::max(1u, p.capacity))))
        return rc;
    HIPCHK(hipMemset(p.count, 0, sizeof(unsigned)));
    h->probes.push_back(p);
    h->probes_dirty = true;
    if (id) *id = (int)h->probes.size() - 1;
    return 0;
}

int ems_probe_results(ems_engine* h, int id, unsigned* count, unsigned* ts, double* values)
{
    if (!h || id < 0 || id >= (int)h->probes.size() || !count) return fail(EMS_E_ARG, "bad probe id");
    HIPCHK(hipSetDevice(h->device));
    HIPCHK(hipStreamSynchronize(h->stream));
    const ProbeDev& p = h->probes[id];
    unsigned c = 0;
    HIPCHK(hipMemcpy(&c, p.count, sizeof(unsigned), hipMemcpyDeviceToHost));
    *count = c;
    if (ts && c) HIPCHK(hipMemcpy(ts, p.ts_out, c * sizeof(unsigned), hipMemcpyDeviceToHost));
    if (values && c) HIPCHK(hipMemcpy(values, p.val_out, 3 * (size_t)c * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

}  // extern "C"

// box, field, interpolation and edge lengths of a new dump
static int dump_common(ems_engine* h, int field, int interp, const unsigned start[3], const unsigned stop[3],
                       const double* const edge_len[3], DumpDev& d)
{
    if (!h || !start || !stop || !edge_len) return fail(EMS_E_ARG, "null argument");
    if (field != 0 && field != 1) return fail(EMS_E_ARG, "dump field must be 0 (E) or 1 (H)");
    if (interp < 0 || interp > 2) return fail(EMS_E_ARG, "dump interpolation must be 0 (none), 1 (node) or 2 (cell)");
    for (int n = 0; n < 3; ++n) {
        if (start[n] > stop[n] || stop[n] >= h->N[n]) return fail(EMS_E_ARG, "dump box outside the mesh");
        d.s[n] = start[n];
        d.n[n] = stop[n] - start[n] + 1;
        d.N[n] = h->N[n];
    }
    d.field = field;
    d.interp = interp;
    HIPCHK(hipSetDevice(h->device));
    for (int n = 0; n < 3; ++n) {
        double* de;
        int rc;
        if ((rc = upload(h, &de, edge_len[n], h->N[n]))) return rc;
        d.el[n] = de;
    }
    return 0;
}

extern "C" {

int ems_add_fd_dump(ems_engine* h, int field, int interp, const unsigned start[3], const unsigned stop[3],
                    unsigned fd_interval, unsigned first_ts, unsigned nfreq, unsigned nsamples, const float* factors,
                    const double* const edge_len[3], int* id)
{
    if (!h || !factors) return fail(EMS_E_ARG, "null argument");
    if (fd_interval == 0 || nfreq == 0 || nsamples == 0) return fail(EMS_E_ARG, "empty frequency-domain dump");
    if (first_ts % fd_interval) return fail(EMS_E_ARG, "first dump sample not on the sampling interval");
    DumpDev d{};
    int rc;
    if ((rc = dump_common(h, field, interp, start, stop, edge_len, d))) return rc;
    drop_graphs(h);     // the dump changes the launch pattern of every chunk
    d.fd_interval = fd_interval;
    d.first = first_ts;
    d.nfreq = nfreq;
    d.nsamples = nsamples;
    float* df;
    if ((rc = upload(h, &df, factors, (size_t)nsamples * nfreq * 2))) return rc;
    d.factors = df;
    const size_t npts = (size_t)d.n[0] * d.n[1] * d.n[2];
    if ((rc = h->alloc(&d.acc, (size_t)nfreq * 3 * npts * 2))) return rc;
    HIPCHK(hipMemset(d.acc, 0, (size_t)nfreq * 3 * npts * 2 * sizeof(float)));
    h->dumps.push_back(d);
    h->dump_npts.push_back((unsigned)npts);
    if (id) *id = (int)h->dumps.size() - 1;
    return 0;
}

int ems_fd_dump_results(ems_engine* h, int id, float* out)
{
    if (!h || id < 0 || id >= (int)h->dumps.size() || !out) return fail(EMS_E_ARG, "bad dump id");
    const DumpDev& d = h->dumps[id];
    if (d.td) return fail(EMS_E_ARG, "dump id is a time-domain dump");
    HIPCHK(hipSetDevice(h->device));
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipMemcpy(out, d.acc, (size_t)d.nfreq * 3 * h->dump_npts[id] * 2 * sizeof(float), hipMemcpyDeviceToHost));
    return 0;
}

int ems_add_td_dump(ems_engine* h, int field, int interp, const unsigned start[3], const unsigned stop[3],
                    unsigned interval, unsigned start_ts, unsigned stop_ts, unsigned capacity,
                    const double* const edge_len[3], int* id)
{
    if (!h) return fail(EMS_E_ARG, "null argument");
    if (interval == 0 || capacity == 0) return fail(EMS_E_ARG, "empty time-domain dump");
    if (stop_ts < start_ts) return fail(EMS_E_ARG, "dump window ends before it starts");
    DumpDev d{};
    int rc;
    if ((rc = dump_common(h, field, interp, start, stop, edge_len, d))) return rc;
    const unsigned long long first = ((unsigned long long)start_ts + interval - 1) / interval * interval;
    if (first > stop_ts) return fail(EMS_E_ARG, "no sampling step inside the dump window");
    const unsigned long long inwin = ((unsigned long long)stop_ts - first) / interval + 1;
    drop_graphs(h);
    d.td = 1;
    d.fd_interval = interval;
    d.first = (unsigned)first;
    d.nsamples = (unsigned)std::min<unsigned long long>(capacity, inwin);
    const size_t npts = (size_t)d.n[0] * d.n[1] * d.n[2];
    if ((rc = h->alloc(&d.acc, (size_t)d.nsamples * 3 * npts))) return rc;
    HIPCHK(hipMemset(d.acc, 0, (size_t)d.nsamples * 3 * npts * sizeof(float)));
    h->dumps.push_back(d);
    h->dump_npts.push_back((unsigned)npts);
    if (id) *id = (int)h->dumps.size() - 1;
    return 0;
}

int ems_td_dump_results(ems_engine* h, int id, unsigned* count, unsigned* ts, float* out)
{
    if (!h || id < 0 || id >= (int)h->dumps.size() || !count) return fail(EMS_E_ARG, "bad dump id");
    const DumpDev& d = h->dumps[id];
    if (!d.td) return fail(EMS_E_ARG, "dump id is a frequency-domain dump");
    unsigned c = 0;
    if (h->numTS >= d.first && (h->numTS > 0 || h->sampled0))
        c = (unsigned)std::min<unsigned long long>(d.nsamples, (h->numTS - d.first) / d.fd_interval + 1);
    *count = c;
    if (ts)
        for (unsigned k = 0; k < c; ++k) ts[k] = d.first + k * d.fd_interval;
    if (out && c) {
        HIPCHK(hipSetDevice(h->device));
        HIPCHK(hipStreamSynchronize(h->stream));
        HIPCHK(hipMemcpy(out, d.acc, (size_t)c * 3 * h->dump_npts[id] * sizeof(float), hipMemcpyDeviceToHost));
    }
    return 0;
}

static int sync_probe_tables(ems_engine* h)
{
    if (!h->probes_dirty) return 0;
    // captured chunks hold the old probe tables
    drop_graphs(h);
    int rc;
    if ((rc = upload(h, &h->d_probes, h->probes.data(), h->probes.size())) ||
        (rc = upload(h, &h->d_elin, h->h_elin.data(), h->h_elin.size())) ||
        (rc = upload(h, &h->d_esign, h->h_esign.data(), h->h_esign.size())) ||
        (rc = upload(h, &h->d_ediv, h->h_ediv.data(), h->h_ediv.size())))
        return rc;
    h->probes_dirty = false;
    return 0;
}

static bool probe_due(const ems_engine* h, unsigned ts)
{
    for (const auto& p : h->probes) {
        if (ts < p.start_ts || ts > p.stop_ts) continue;
        if ((p.interval && ts % p.interval == 0) || (p.fd_interval && ts % p.fd_interval == 0)) return true;
    }
    return false;
}

static int launch_samples(ems_engine* h)
{
    // ProcessingArray::Process after the step: probes and dumps due at numTS
    if (!h->probes.empty() && probe_due(h, h->numTS)) {
        hipLaunchKernelGGL(k_probe, dim3((unsigned)h->probes.size()), dim3(kBlock), 0, h->stream, h->V, h->I,
                           h->d_probes, h->d_elin, h->d_esign, h->d_ediv, h->numts_dev);
        HIPCHK(hipGetLastError());
    }
    for (size_t k = 0; k < h->dumps.size(); ++k) {
        const DumpDev& d = h->dumps[k];
        unsigned ks;
        if (!dump_sample(d, h->numTS, &ks)) continue;
        const unsigned npts = h->dump_npts[k];
        hipLaunchKernelGGL(k_fd_dump, dim3((npts + kBlock - 1) / kBlock), dim3(kBlock), 0, h->stream,
                           d.field ? h->I : h->V, h->cs, h->N[0], h->g.plane, d, h->numts_dev);
        HIPCHK(hipGetLastError());
    }
    return 0;
}

}  // extern "C"

namespace {

template <int PM, bool CC>
void launch_volt_t(ems_engine* h, unsigned nblk)
{
    if (h->nt_fields)
        hipLaunchKernelGGL((k_volt<PM, CC, true>), dim3(nblk), dim3(kBlock), 0, h->stream, h->g, h->V, h->I, h->kV, h->pmlV,
                           h->qsrc, h->numts_dev);
    else
        hipLaunchKernelGGL((k_volt<PM, CC, false>), dim3(nblk), dim3(kBlock), 0, h->stream, h->g, h->V, h->I, h->kV,
                           h->pmlV, h->qsrc, h->numts_dev);
}

template <int PM, bool CC>
void launch_curr_t(ems_engine* h, unsigned nblk)
{
    if (h->nt_fields)
        hipLaunchKernelGGL((k_curr<PM, CC, true>), dim3(nblk), dim3(kBlock), 0, h->stream, h->g, h->I, h->V, h->kI,
                           h->pmlI, h->numts_dev);
    else
        hipLaunchKernelGGL((k_curr<PM, CC, false>), dim3(nblk), dim3(kBlock), 0, h->stream, h->g, h->I, h->V, h->kI,
                           h->pmlI, h->numts_dev);
}

template <bool CC>
void launch_update(ems_engine* h, bool curr, unsigned nblk)
{
    const int pm = h->pmlV.n ? h->pml_mode : kNoPml;
    if (pm == kPmlCompressed) curr ? launch_curr_t<kPmlCompressed, CC>(h, nblk) : launch_volt_t<kPmlCompressed, CC>(h, nblk);
    else if (pm == kPmlPlain) curr ? launch_curr_t<kPmlPlain, CC>(h, nblk) : launch_volt_t<kPmlPlain, CC>(h, nblk);
    else curr ? launch_curr_t<kNoPml, CC>(h, nblk) : launch_volt_t<kNoPml, CC>(h, nblk);
}

template <int PM, bool CC, bool LT>
void launch_quad_box(ems_engine* h, bool curr, const QBoxes& B, unsigned nblk, unsigned* numts)
{
    if (!nblk) return;
    const size_t lds = LT ? (size_t)h->kV.ntab * sizeof(float4) : 0;
    if (curr) {
        if (h->nt_fields)
            hipLaunchKernelGGL((k_curr4<PM, CC, true, LT>), dim3(nblk), dim3(kBlock), lds, h->stream, h->gq, B, h->I,
                               h->V, h->kI, h->pmlI, numts);
        else
            hipLaunchKernelGGL((k_curr4<PM, CC, false, LT>), dim3(nblk), dim3(kBlock), lds, h->stream, h->gq, B, h->I,
                               h->V, h->kI, h->pmlI, numts);
    } else {
        if (h->nt_fields)
            hipLaunchKernelGGL((k_volt4<PM, CC, true, LT>), dim3(nblk), dim3(kBlock), lds, h->stream, h->gq, B, h->V,
                               h->I, h->kV, h->pmlV, h->qsrc, h->numts_dev);
        else
            hipLaunchKernelGGL((k_volt4<PM, CC, false, LT>), dim3(nblk), dim3(kBlock), lds, h->stream, h->gq, B, h->V,
                               h->I, h->kV, h->pmlV, h->qsrc, h->numts_dev);
    }
}

template <int PM, bool CC, bool LT>
void launch_quad_lt(ems_engine* h, bool curr)
{
    if (PM != kNoPml && h->qsplit) {
        // the box without UPML cells runs without UPML code; the boxes around
        // it (and the timestep counter) take the UPML path
        launch_quad_box<kNoPml, CC, LT>(h, curr, h->qbox_in, h->qblk_in, nullptr);
        launch_quad_box<PM, CC, LT>(h, curr, h->qbox_out, h->qblk_out, h->numts_dev);
        return;
    }
    launch_quad_box<PM, CC, LT>(h, curr, h->qbox_all, h->qblk_all, h->numts_dev);
}

template <int PM, bool CC>
void launch_quad_t(ems_engine* h, bool curr)
{
    if (CC && h->lds_tab) launch_quad_lt<PM, CC, true>(h, curr);
    else launch_quad_lt<PM, CC, false>(h, curr);
}

template <bool CC>
void launch_quad(ems_engine* h, bool curr)
{
    const int pm = h->pmlV.n ? h->pml_mode : kNoPml;
    if (pm == kPmlCompressed) launch_quad_t<kPmlCompressed, CC>(h, curr);
    else if (pm == kPmlPlain) launch_quad_t<kPmlPlain, CC>(h, curr);
    else launch_quad_t<kNoPml, CC>(h, curr);
}

void dispatch_update(ems_engine* h, bool curr, unsigned nblk)
{
    if (h->quad) {
        if (h->cmp) launch_quad<true>(h, curr);
        else launch_quad<false>(h, curr);
        return;
    }
    if (h->cmp) launch_update<true>(h, curr, nblk);
    else launch_update<false>(h, curr, nblk);
}

template <int PM, bool CC>
void launch_step_t(ems_engine* h, const StepArgs& a, unsigned nblk)
{
    hipLaunchKernelGGL((k_step<PM, CC>), dim3(nblk), dim3(kBlock), 0, h->stream, a);
}

void lds_step(ems_engine* h);

}  // namespace

// state n+1 written by a single-sweep step becomes the current state
static void swap_step_buffers(ems_engine* h)
{
    std::swap(h->V, h->V_alt);
    std::swap(h->I, h->I_alt);
    for (int side = 0; side < 2; ++side) {
        PmlArgs& P = side ? h->pmlI : h->pmlV;
        if (!P.n) continue;
        float* cur = const_cast<float*>(P.flux_in);
        P.flux_in = h->flux_alt[side];
        P.flux_out = h->flux_alt[side];
        h->flux_alt[side] = cur;
    }
}

namespace {

template <int PM, bool CC>
void launch_step2_t(ems_engine* h, const Step2Args& a, unsigned nblk)
{
    hipLaunchKernelGGL((k_step2<PM, CC>), dim3(nblk), dim3(kBlock), 0, h->stream, a, h->V, h->I, h->V_alt, h->I_alt,
                       h->pmlV.flux_in, h->flux_alt[0], h->pmlI.flux_in, h->flux_alt[1]);
}

void column_step(ems_engine* h)
{
    Step2Args a{};
    a.g = h->g;
    a.kV = h->kV;
    a.kI = h->kI;
    a.pV = h->pmlV;
    a.pI = h->pmlI;
    a.eV = h->eV;
    a.eI = h->eI;
    a.numts = h->numts_dev;
    const unsigned nblk = h->g.qtiles * h->g.nzc;
    const int pm = h->pmlV.n ? h->pml_mode : kNoPml;
    if (h->cmp) {
        if (pm == kPmlCompressed) launch_step2_t<kPmlCompressed, true>(h, a, nblk);
        else if (pm == kPmlPlain) launch_step2_t<kPmlPlain, true>(h, a, nblk);
        else launch_step2_t<kNoPml, true>(h, a, nblk);
    } else {
        if (pm == kPmlCompressed) launch_step2_t<kPmlCompressed, false>(h, a, nblk);
        else if (pm == kPmlPlain) launch_step2_t<kPmlPlain, false>(h, a, nblk);
        else launch_step2_t<kNoPml, false>(h, a, nblk);
    }
}

template <bool CC, bool LT>
void launch_step4_lt(ems_engine* h, const Step4Args& a, unsigned nblk)
{
    const size_t lds = LT ? (size_t)(h->kV.ntab + h->kI.ntab) * sizeof(float4) : 0;
    if (h->q4tile) {
        if (h->nt_fields)
            hipLaunchKernelGGL((k_step4t<CC, true, LT>), dim3(nblk), dim3(kBlock), lds, h->stream, a, h->V, h->I,
                               h->V_alt, h->I_alt);
        else
            hipLaunchKernelGGL((k_step4t<CC, false, LT>), dim3(nblk), dim3(kBlock), lds, h->stream, a, h->V, h->I,
                               h->V_alt, h->I_alt);
        return;
    }
    if (h->nt_fields)
        hipLaunchKernelGGL((k_step4<CC, true, LT>), dim3(nblk), dim3(kBlock), lds, h->stream, a, h->V, h->I, h->V_alt,
                           h->I_alt);
    else
        hipLaunchKernelGGL((k_step4<CC, false, LT>), dim3(nblk), dim3(kBlock), lds, h->stream, a, h->V, h->I, h->V_alt,
                           h->I_alt);
}

template <bool CC>
void launch_step4_t(ems_engine* h, const Step4Args& a, unsigned nblk)
{
    if (CC && h->lds_tab) launch_step4_lt<CC, true>(h, a, nblk);
    else launch_step4_lt<CC, false>(h, a, nblk);
}

// the quad column sweep over the whole mesh (no UPML)
void quad_step(ems_engine* h)
{
    Step4Args a{};
    a.g = h->g;
    a.kV = h->kV;
    a.kI = h->kI;
    a.eV = h->eV;
    a.eI = h->eI;
    a.numts = h->numts_dev;
    a.qa = 0;
    a.nq = h->N[0] / 4;
    a.ya = 0;
    a.ny = h->N[1];
    a.za = 0;
    a.zb = h->N[2];
    a.xb = h->N[0];
    a.yb = h->N[1];
    a.zdepth = h->fzdepth;
    a.ntiles = h->ntx;
    a.ntq = h->q4ntq;
    for (int side = 0; side < 2; ++side) {
        if (h->mur_fold[side] < 0) continue;
        const MurDev& m = h->mur[h->mur_fold[side]];
        a.mf[side] = MurFold{m.cP, m.cPP, m.mP, m.mPP, m.start_ts, 1u};
    }
    const unsigned nblk = h->ntx * h->fnzc;
    if (h->cmp) launch_step4_t<true>(h, a, nblk);
    else launch_step4_t<false>(h, a, nblk);
}

// one fused step from state n (h->V, h->I, current fluxes) into the second
// buffers, which then become the current state
void fused_step(ems_engine* h, hipEvent_t* ev)
{
    const unsigned T = 256;
    // Mur pre hooks read state n (FDTD/engine.cpp:224-265 order; only the quad sweep has Mur)
    for (size_t k = 0; k < h->mur.size(); ++k) {
        if ((int)k == h->mur_fold[0] || (int)k == h->mur_fold[1]) continue;     // in k_step4t
        const auto& m = h->mur[k];
        hipLaunchKernelGGL(k_mur_pre, dim3((m.na * m.nb + T - 1) / T), dim3(T), 0, h->stream, h->V, m, h->numts_dev);
    }
    if (ev) (void)hipEventRecord(ev[0], h->stream);
    if (h->fused == 3) quad_step(h);
    else if (h->fused == 2) column_step(h);
    else lds_step(h);
    if (ev) {
        (void)hipEventRecord(ev[1], h->stream);
        (void)hipEventRecord(ev[2], h->stream);
    }
    if (!h->mur.empty()) {
        // state n+1: Mur post (reverse creation order) on the swept voltages,
        // Mur apply, the deferred voltage sources, then the current update of
        // the planes next to each face again, with their current sources
        for (int k = (int)h->mur.size() - 1; k >= 0; --k) {
            if (k == h->mur_fold[0] || k == h->mur_fold[1]) continue;          // in k_step4t
            const auto& m = h->mur[k];
            hipLaunchKernelGGL(k_mur_post, dim3((m.na * m.nb + T - 1) / T), dim3(T), 0, h->stream, h->V_alt, m,
                               h->numts_dev);
        }
        for (int k = (int)h->mur.size() - 1; k >= 0; --k) {
            const auto& m = h->mur[k];
            const int side = k == h->mur_fold[0] ? 0 : k == h->mur_fold[1] ? 1 : -1;
            if (side >= 0) {
                if (h->nrim[side])
                    hipLaunchKernelGGL(k_mur_apply_rim, dim3((h->nrim[side] + T - 1) / T), dim3(T), 0, h->stream,
                                       h->V_alt, m, h->mur_rim[side], h->nrim[side], h->numts_dev);
                continue;
            }
            hipLaunchKernelGGL(k_mur_apply, dim3((m.na * m.nb + T - 1) / T), dim3(T), 0, h->stream, h->V_alt, m,
                               h->numts_dev);
        }
        const auto& e0 = h->exc_def[0];
        if (e0.G)
            hipLaunchKernelGGL(k_excite, dim3((e0.G + T - 1) / T), dim3(T), 0, h->stream, h->V_alt, e0.G, e0.gbeg,
                               e0.glin, e0.amp, e0.delay, h->sig[0], h->sigLen, h->period, h->numts_dev, 0);
        for (const auto& b : h->bands) {
            const unsigned long long cnt = (unsigned long long)(b.hi - b.lo + 1) * h->N[0] * h->N[1] * h->N[2] /
                                           h->N[b.axis];
            hipLaunchKernelGGL(k_curr_band, dim3((unsigned)((cnt + T - 1) / T)), dim3(T), 0, h->stream, h->g,
                               h->V_alt, h->I, h->I_alt, h->kI, b.axis, b.lo, b.hi);
        }
        if (h->ncells)
            hipLaunchKernelGGL(k_curr_cells, dim3((h->ncells + T - 1) / T), dim3(T), 0, h->stream, h->g, h->V_alt,
                               h->I, h->I_alt, h->kI, h->cell_list, h->ncells);
        const auto& e1 = h->exc_def[1];
        if (e1.G)
            hipLaunchKernelGGL(k_excite, dim3((e1.G + T - 1) / T), dim3(T), 0, h->stream, h->I_alt, e1.G, e1.gbeg,
                               e1.glin, e1.amp, e1.delay, h->sig[1], h->sigLen, h->period, h->numts_dev, 0);
    }
    hipLaunchKernelGGL(k_tick, dim3(1), dim3(64), 0, h->stream, h->numts_dev);
    if (ev) (void)hipEventRecord(ev[3], h->stream);
    swap_step_buffers(h);
}

void lds_step(ems_engine* h)
{
    StepArgs a{};
    a.g = h->g;
    a.g.zdepth = h->fzdepth;
    a.g.nzc = h->fnzc;
    a.ntx = h->ntx;
    a.nty = h->nty;
    a.V0 = h->V;
    a.I0 = h->I;
    a.V1 = h->V_alt;
    a.I1 = h->I_alt;
    a.kV = h->kV;
    a.kI = h->kI;
    a.pV = h->pmlV;
    a.pI = h->pmlI;
    if (a.pV.n) a.pV.flux_out = h->flux_alt[0];
    if (a.pI.n) a.pI.flux_out = h->flux_alt[1];
    a.eV = h->eV;
    a.eI = h->eI;
    a.numts = h->numts_dev;
    const unsigned nblk = h->ntx * h->nty * h->fnzc;
    const int pm = h->pmlV.n ? h->pml_mode : kNoPml;
    if (h->cmp) {
        if (pm == kPmlCompressed) launch_step_t<kPmlCompressed, true>(h, a, nblk);
        else if (pm == kPmlPlain) launch_step_t<kPmlPlain, true>(h, a, nblk);
        else launch_step_t<kNoPml, true>(h, a, nblk);
    } else {
        if (pm == kPmlCompressed) launch_step_t<kPmlCompressed, false>(h, a, nblk);
        else if (pm == kPmlPlain) launch_step_t<kPmlPlain, false>(h, a, nblk);
        else launch_step_t<kNoPml, false>(h, a, nblk);
    }
}

}  // namespace

extern "C" {

// one timestep in the reference's hook order (Engine::IterateTS, FDTD/engine.cpp:267-286,
// priorities: UPML > default (lumped RLC, Lorentz, Mur) > excitation)
static int launch_step(ems_engine* h, hipEvent_t* ev)
{
    if (h->fused) {
        fused_step(h, ev);
        HIPCHK(hipGetLastError());
        return 0;
    }
    const GridDesc& g = h->g;
    const unsigned nblk = g.qtiles * g.nzc;
    const unsigned T = 256;
    // -- voltage pre hooks: Mur (creation order), Lorentz, lumped RLC; UPML pre is fused
    for (size_t k = 0; k < h->mur.size(); ++k) {
        if ((int)k == h->mur_fold[0] || (int)k == h->mur_fold[1]) continue;     // in k_step4t
        const auto& m = h->mur[k];
        hipLaunchKernelGGL(k_mur_pre, dim3((m.na * m.nb + T - 1) / T), dim3(T), 0, h->stream, h->V, m, h->numts_dev);
    }
    for (size_t k = 0; k < h->lor.size(); k += 2)
        if (h->lor_v[k / 2])
            hipLaunchKernelGGL(k_lor_pre, dim3((h->lor[k].count + T - 1) / T), dim3(T), 0, h->stream, h->V, h->cs, h->lor[k]);
    for (const auto& r : h->rlc)
        hipLaunchKernelGGL(k_rlc_pre, dim3((r.count + T - 1) / T), dim3(T), 0, h->stream, r, h->numts_dev);
    if (ev) (void)hipEventRecord(ev[0], h->stream);
    dispatch_update(h, false, nblk);
    if (ev) (void)hipEventRecord(ev[1], h->stream);
    // -- post hooks: Mur (reverse creation order, all read-only on V)
    for (int k = (int)h->mur.size() - 1; k >= 0; --k) {
        const auto& m = h->mur[k];
        hipLaunchKernelGGL(k_mur_post, dim3((m.na * m.nb + T - 1) / T), dim3(T), 0, h->stream, h->V, m, h->numts_dev);
    }
    // -- apply hooks: steady-state record (priority +2e6), lumped RLC, Lorentz,
    //    Mur (reverse creation order), excitation
    if (h->has_ss) {
        const SteadyDev& S = h->ss;
        hipLaunchKernelGGL(k_ss_record, dim3((S.n + T - 1) / T), dim3(T), 0, h->stream, h->V, S, h->numts_dev);
        if (h->numTS % S.p == 0 && h->numTS >= 2 * S.p) {
            const unsigned nb = 1024;
            hipLaunchKernelGGL(k_energy, dim3(nb), dim3(kBlock), 0, h->stream, h->V, h->I, h->cs, h->N[0], h->N[1],
                               h->N[2], h->ss_part);
            hipLaunchKernelGGL(k_ss_eval, dim3(1), dim3(64), 0, h->stream, S, h->ss_part, nb, h->numts_dev);
        }
    }
    for (const auto& r : h->rlc)
        hipLaunchKernelGGL(k_rlc_apply, dim3((r.count + T - 1) / T), dim3(T), 0, h->stream, h->V, r, h->numts_dev);
    for (size_t k = 0; k < h->lor.size(); k += 2)
        if (h->lor_v[k / 2])
            hipLaunchKernelGGL(k_lor_apply, dim3((h->lor[k].count + T - 1) / T), dim3(T), 0, h->stream, h->V, h->cs, h->lor[k]);
    for (int k = (int)h->mur.size() - 1; k >= 0; --k) {
        const auto& m = h->mur[k];
        hipLaunchKernelGGL(k_mur_apply, dim3((m.na * m.nb + T - 1) / T), dim3(T), 0, h->stream, h->V, m, h->numts_dev);
    }
    if (h->exc[0].G && !h->qsrc.beg) {
        const auto& e = h->exc[0];
        hipLaunchKernelGGL(k_excite, dim3((e.G + T - 1) / T), dim3(T), 0, h->stream, h->V, e.G, e.gbeg, e.glin, e.amp,
                           e.delay, h->sig[0], h->sigLen, h->period, h->numts_dev, 0);
    }
    // -- current half step: Lorentz pre, fused UPML + update, Lorentz apply, excitation
    for (size_t k = 1; k < h->lor.size(); k += 2)
        if (h->lor_i[k / 2])
            hipLaunchKernelGGL(k_lor_pre, dim3((h->lor[k].count + T - 1) / T), dim3(T), 0, h->stream, h->I, h->cs, h->lor[k]);
    if (ev) (void)hipEventRecord(ev[2], h->stream);
    dispatch_update(h, true, nblk);
    if (ev) (void)hipEventRecord(ev[3], h->stream);
    for (size_t k = 1; k < h->lor.size(); k += 2)
        if (h->lor_i[k / 2])
            hipLaunchKernelGGL(k_lor_apply, dim3((h->lor[k].count + T - 1) / T), dim3(T), 0, h->stream, h->I, h->cs, h->lor[k]);
    if (h->exc[1].G) {
        const auto& e = h->exc[1];
        hipLaunchKernelGGL(k_excite, dim3((e.G + T - 1) / T), dim3(T), 0, h->stream, h->I, e.G, e.gbeg, e.glin, e.amp,
                           e.delay, h->sig[1], h->sigLen, h->period, h->numts_dev, -1);
    }
    HIPCHK(hipGetLastError());
    return 0;
}

int ems_set_profiling(ems_engine* h, int enable)
{
    if (!h) return fail(EMS_E_ARG, "null handle");
    h->profile = enable != 0;
    return 0;
}

// steps of one ems_iterate call whose kernels are bracketed by events
constexpr unsigned kProfSteps = 4096;

// Steps per captured chunk: the requested length rounded up to a multiple of
// the period of the launch pattern (the sampling intervals of probes and
// dumps, the steady-state period, 2 for the buffer parity of a single-sweep
// form) when that period is at most 1024 steps, so that chunks repeat.
static unsigned graph_steps(const ems_engine* h)
{
    unsigned long long P = h->fused ? 2 : 1;
    auto add = [&P](unsigned v) {
        if (v) P = P / std::gcd(P, (unsigned long long)v) * v;
    };
    for (const auto& p : h->probes) {
        add(p.interval);
        add(p.fd_interval);
    }
    for (const auto& d : h->dumps) add(d.fd_interval);
    if (h->has_ss) add(h->ss.p);
    const unsigned c = h->graph_chunk;
    if (P > 1024) return c;
    return (unsigned)((c + P - 1) / P * P);
}

// launch pattern of the next `steps` steps from the host timestep on
static std::vector<unsigned char> chunk_key(const ems_engine* h, unsigned steps)
{
    std::vector<unsigned char> key;
    key.push_back((unsigned char)(h->fused && h->V != h->V_base ? 1 : 0));
    for (unsigned k = 0; k < steps; ++k) {
        const unsigned ts = h->numTS + k;          // numTS before step k of the chunk
        unsigned char b = 0;
        if (h->has_ss && ts % h->ss.p == 0 && ts >= 2 * h->ss.p) b |= 1;
        if (!h->probes.empty() && probe_due(h, ts + 1)) b |= 2;
        key.push_back(b);
        for (size_t d = 0; d < h->dumps.size(); ++d) {
            unsigned ks;
            key.push_back((unsigned char)dump_sample(h->dumps[d], ts + 1, &ks));
        }
    }
    return key;
}

// one chunk of graph_chunk steps: replay the graph of its launch pattern,
// captured from the direct launches the first time the pattern occurs
static int graph_chunk(ems_engine* h, unsigned steps)
{
    // liveness of what the captured launches reference: every buffer a graph
    // names is an allocation of this engine, kept until ems_destroy; one made
    // after a capture (a table uploaded again) means the graphs may name the
    // previous one, so they are captured anew
    if (!h->graphs.empty() && h->graphs.front().nalloc != h->allocs.size()) drop_graphs(h);
    const std::vector<unsigned char> key = chunk_key(h, steps);
    ems_engine::GraphEnt* G = nullptr;
    for (auto& e : h->graphs)
        if (e.key == key) G = &e;
    if (G) {
        HIPCHK(hipGraphLaunch(G->exec, h->stream));
        // host state as the captured launches left it
        for (unsigned k = 0; k < steps; ++k) {
            if (h->fused) swap_step_buffers(h);
            ++h->numTS;
        }
        ++h->graph_replays;
        return 0;
    }
    if (h->graphs.size() >= kMaxGraphs) {
        // too many distinct patterns: direct launches
        for (unsigned k = 0; k < steps; ++k) {
            int rc = launch_step(h, nullptr);
            ++h->numTS;
            if (!rc) rc = launch_samples(h);
            if (rc) return rc;
        }
        return 0;
    }
    ems_engine::GraphEnt e;
    e.key = key;
    HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    int lrc = 0;
    for (unsigned k = 0; k < steps; ++k) {
        if (!lrc) lrc = launch_step(h, nullptr);
        ++h->numTS;
        if (!lrc) lrc = launch_samples(h);
    }
    const hipError_t ec = hipStreamEndCapture(h->stream, &e.graph);
    if (lrc) {
        if (ec == hipSuccess && e.graph) (void)hipGraphDestroy(e.graph);
        return lrc;
    }
    if (ec != hipSuccess) return fail(EMS_E_HIP, std::string("graph capture failed: ") + hipGetErrorString(ec));
    HIPCHK(hipGraphInstantiate(&e.exec, e.graph, nullptr, nullptr, 0));
    HIPCHK(hipGraphLaunch(e.exec, h->stream));
    e.nalloc = h->allocs.size();
    h->graphs.push_back(std::move(e));
    ++h->graph_captures;
    return 0;
}

static void drop_graphs(ems_engine* h)
{
    if (!h->graphs.empty()) hipStreamSynchronize(h->stream);
    for (auto& G : h->graphs) {
        if (G.exec) hipGraphExecDestroy(G.exec);
        if (G.graph) hipGraphDestroy(G.graph);
    }
    h->graphs.clear();
}

int ems_set_graph(ems_engine* h, unsigned chunk_steps)
{
    if (!h) return fail(EMS_E_ARG, "null handle");
    if (chunk_steps > 4096) return fail(EMS_E_ARG, "graph chunk above 4096 steps");
    drop_graphs(h);
    h->graph_chunk = chunk_steps;
    return 0;
}

int ems_get_graph(const ems_engine* h, unsigned* chunk_steps, unsigned* captures, unsigned* replays)
{
    if (!h) return fail(EMS_E_ARG, "null handle");
    if (chunk_steps) *chunk_steps = h->graph_chunk;
    if (captures) *captures = h->graph_captures;
    if (replays) *replays = h->graph_replays;
    return 0;
}

int ems_iterate(ems_engine* h, unsigned nsteps)
{
    if (!h) return fail(EMS_E_ARG, "null handle");
    TxRange tx("ems_iterate");
    int rc;
    if (!h->inited && (rc = ems_init(h))) return rc;
    HIPCHK(hipSetDevice(h->device));
    rc = sync_probe_tables(h);
    if (rc) return rc;
    const unsigned nprof = h->profile ? std::min(nsteps, kProfSteps) : 0;
    while (h->kev.size() < 4 * (size_t)nprof) {
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        h->kev.push_back(e);
    }
    if (h->numTS == 0 && nsteps > 0) {
        // Process() before the first step (openems.cpp:1424-1425)
        if ((rc = launch_samples(h))) return rc;
        h->sampled0 = true;
    }
    HIPCHK(hipEventRecord(h->ev0, h->stream));
    unsigned s = 0;
    for (; s < nprof; ++s) {
        rc = launch_step(h, &h->kev[4 * (size_t)s]);
        ++h->numTS;
        if (!rc) rc = launch_samples(h);
        if (rc) return rc;
    }
    if (h->graph_chunk) {
        const unsigned steps = graph_steps(h);
        if (steps != h->graph_len) {
            drop_graphs(h);
            h->graph_len = steps;
        }
        while (nsteps - s >= steps) {
            TxRange txc("ems_iterate chunk");
            if ((rc = graph_chunk(h, steps))) return rc;
            s += steps;
        }
    }
    for (; s < nsteps; ++s) {
        rc = launch_step(h, nullptr);
        ++h->numTS;
        if (!rc) rc = launch_samples(h);
        if (rc) return rc;
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(h->ev1, h->stream));
    if (!h->async || nprof) {
        HIPCHK(hipEventSynchronize(h->ev1));
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, h->ev0, h->ev1));
        h->last_ms = ms;
    }
    if (nprof) {
        double vsum = 0, csum = 0;
        for (unsigned s = 0; s < nprof; ++s) {
            float a = 0, b = 0;
            HIPCHK(hipEventElapsedTime(&a, h->kev[4 * (size_t)s], h->kev[4 * (size_t)s + 1]));
            HIPCHK(hipEventElapsedTime(&b, h->kev[4 * (size_t)s + 2], h->kev[4 * (size_t)s + 3]));
            vsum += a;
            csum += b;
        }
        h->last_volt_ms = vsum / nprof;
        h->last_curr_ms = csum / nprof;
    }
    return 0;
}

unsigned ems_get_numts(const ems_engine* h) { return h ? h->numTS : 0; }

int ems_set_async(ems_engine* h, int async)
{
    if (!h) return fail(EMS_E_ARG, "null handle");
    h->async = async != 0;
    return 0;
}

int ems_sync(ems_engine* h)
{
    if (!h) return fail(EMS_E_ARG, "null handle");
    HIPCHK(hipSetDevice(h->device));
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

static int get_field(ems_engine* h, const float* d, float* out)
{
    if (!h || !out) return fail(EMS_E_ARG, "null argument");
    HIPCHK(hipSetDevice(h->device));
    HIPCHK(hipStreamSynchronize(h->stream));
    std::vector<float> stage(3 * h->cs);
    HIPCHK(hipMemcpy(stage.data(), d, stage.size() * sizeof(float), hipMemcpyDeviceToHost));
    soa_to_nijk(stage.data(), out, h->N, h->cs);
    return 0;
}

static int set_field(ems_engine* h, float* d, float* alt, const float* in)
{
    if (!h || !in) return fail(EMS_E_ARG, "null argument");
    HIPCHK(hipSetDevice(h->device));
    HIPCHK(hipStreamSynchronize(h->stream));
    int rc = upload_nijk(h, d, in);
    if (rc || !alt) return rc;
    // the fused step's second buffer must agree on the cells a step leaves alone
    HIPCHK(hipMemcpy(alt, d, 3 * h->cs * sizeof(float), hipMemcpyDeviceToDevice));
    HIPCHK(hipDeviceSynchronize());
    return 0;
}

int ems_get_volt(ems_engine* h, float* out) { return h ? get_field(h, h->V, out) : fail(EMS_E_ARG, "null handle"); }
int ems_get_curr(ems_engine* h, float* out) { return h ? get_field(h, h->I, out) : fail(EMS_E_ARG, "null handle"); }
int ems_set_volt(ems_engine* h, const float* in)
{
    return h ? set_field(h, h->V, h->V_alt, in) : fail(EMS_E_ARG, "null handle");
}
int ems_set_curr(ems_engine* h, const float* in)
{
    return h ? set_field(h, h->I, h->I_alt, in) : fail(EMS_E_ARG, "null handle");
}

// one component's box [s, s+n) as x-fastest planes, through pitched copies
// of the device rows (dir: hipMemcpyDeviceToHost or HostToDevice)
static int box_copy(ems_engine* h, float* dev, float* host_xfast, const unsigned s[3], const unsigned n[3],
                    hipMemcpyKind dir)
{
    const size_t Nx = h->N[0], Ny = h->N[1];
    for (unsigned z = 0; z < n[2]; ++z) {
        float* d = dev + s[0] + Nx * (s[1] + Ny * (size_t)(s[2] + z));
        float* q = host_xfast + (size_t)z * n[0] * n[1];
        if (dir == hipMemcpyDeviceToHost)
            HIPCHK(hipMemcpy2D(q, n[0] * sizeof(float), d, Nx * sizeof(float), n[0] * sizeof(float), n[1], dir));
        else
            HIPCHK(hipMemcpy2D(d, Nx * sizeof(float), q, n[0] * sizeof(float), n[0] * sizeof(float), n[1], dir));
    }
    return 0;
}

static int box_check(const ems_engine* h, int comp, const unsigned start[3], const unsigned stop[3], unsigned n[3])
{
    if (comp < -1 || comp > 2) return fail(EMS_E_ARG, "field component must be 0, 1, 2 or -1 (all three)");
    for (int a = 0; a < 3; ++a) {
        if (start[a] > stop[a] || stop[a] >= h->N[a]) return fail(EMS_E_ARG, "field box outside the mesh");
        n[a] = stop[a] - start[a] + 1;
    }
    return 0;
}

int ems_read_field(ems_engine* h, int is_curr, int comp, const unsigned start[3], const unsigned stop[3], float* out)
{
    if (!h || !start || !stop || !out) return fail(EMS_E_ARG, "null argument");
    unsigned n[3];
    int rc = box_check(h, comp, start, stop, n);
    if (rc) return rc;
    HIPCHK(hipSetDevice(h->device));
    HIPCHK(hipStreamSynchronize(h->stream));
    const size_t cnt = (size_t)n[0] * n[1] * n[2];
    std::vector<float> stage(cnt);
    const int c0 = comp < 0 ? 0 : comp, c1 = comp < 0 ? 2 : comp;
    float* F = is_curr ? h->I : h->V;
    for (int c = c0; c <= c1; ++c) {
        if ((rc = box_copy(h, F + (size_t)c * h->cs, stage.data(), start, n, hipMemcpyDeviceToHost))) return rc;
        float* o = out + (size_t)(c - c0) * cnt;
        for (unsigned x = 0; x < n[0]; ++x)
            for (unsigned y = 0; y < n[1]; ++y)
                for (unsigned z = 0; z < n[2]; ++z)
                    o[((size_t)x * n[1] + y) * n[2] + z] = stage[x + (size_t)n[0] * (y + (size_t)n[1] * z)];
    }
    return 0;
}

int ems_write_field(ems_engine* h, int is_curr, int comp, const unsigned start[3], const unsigned stop[3],
                    const float* in)
{
    if (!h || !start || !stop || !in) return fail(EMS_E_ARG, "null argument");
    unsigned n[3];
    int rc = box_check(h, comp, start, stop, n);
    if (rc) return rc;
    HIPCHK(hipSetDevice(h->device));
    HIPCHK(hipStreamSynchronize(h->stream));
    const size_t cnt = (size_t)n[0] * n[1] * n[2];
    std::vector<float> stage(cnt);
    const int c0 = comp < 0 ? 0 : comp, c1 = comp < 0 ? 2 : comp;
    // the fused step's second buffer must agree on the cells a step leaves alone
    float* bufs[2] = {is_curr ? h->I : h->V, is_curr ? h->I_alt : h->V_alt};
    for (int c = c0; c <= c1; ++c) {
        const float* src = in + (size_t)(c - c0) * cnt;
        for (unsigned x = 0; x < n[0]; ++x)
            for (unsigned y = 0; y < n[1]; ++y)
                for (unsigned z = 0; z < n[2]; ++z)
                    stage[x + (size_t)n[0] * (y + (size_t)n[1] * z)] = src[((size_t)x * n[1] + y) * n[2] + z];
        for (float* F : bufs)
            if (F && (rc = box_copy(h, F + (size_t)c * h->cs, stage.data(), start, n, hipMemcpyHostToDevice))) return rc;
    }
    return 0;
}

int ems_get_field_at(ems_engine* h, int field, int interp, const unsigned pos[3], const double* const edge_len[3],
                     double out[3])
{
    if (!h || !pos || !edge_len || !out) return fail(EMS_E_ARG, "null argument");
    if (field != 0 && field != 1) return fail(EMS_E_ARG, "field must be 0 (E) or 1 (H)");
    if (interp < 0 || interp > 2) return fail(EMS_E_ARG, "interpolation must be 0 (none), 1 (node) or 2 (cell)");
    unsigned lo[3], hi[3], n[3];
    for (int a = 0; a < 3; ++a) {
        if (pos[a] >= h->N[a]) return fail(EMS_E_ARG, "position outside the mesh");
        if (!edge_len[a]) return fail(EMS_E_ARG, "null edge lengths");
        lo[a] = pos[a] ? pos[a] - 1 : 0;
        hi[a] = std::min(pos[a] + 1, h->N[a] - 1);
        n[a] = hi[a] - lo[a] + 1;
    }
    std::vector<float> box(3 * (size_t)n[0] * n[1] * n[2]);
    int rc = ems_read_field(h, field, -1, lo, hi, box.data());
    if (rc) return rc;
    for (int c = 0; c < 3; ++c) {
        const float* C = box.data() + (size_t)c * n[0] * n[1] * n[2];
        auto raw = [&](int dx, int dy, int dz) -> float {
            const unsigned x = pos[0] + dx - lo[0], y = pos[1] + dy - lo[1], z = pos[2] + dz - lo[2];
            return C[((size_t)x * n[1] + y) * n[2] + z];
        };
        out[c] = interp_value(field, interp, h->N, edge_len[c], pos, c, raw);
    }
    return 0;
}

int ems_energy(ems_engine* h, double* out)
{
    if (!h || !out) return fail(EMS_E_ARG, "null argument");
    HIPCHK(hipSetDevice(h->device));
    const unsigned nb = kEnergyBlocks;
    if (!h->energy_part) {
        int rc = h->alloc(&h->energy_part, 2 * nb);
        if (rc) return rc;
    }
    double* part = h->energy_part;
    hipLaunchKernelGGL(k_energy, dim3(nb), dim3(kBlock), 0, h->stream, h->V, h->I, h->cs, h->N[0], h->N[1], h->N[2], part);
    std::vector<double> hp(2 * nb);
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipMemcpy(hp.data(), part, hp.size() * sizeof(double), hipMemcpyDeviceToHost));
    double e = 0, m = 0;
    for (unsigned b = 0; b < nb; ++b) {
        e += hp[2 * b];
        m += hp[2 * b + 1];
    }
    *out = 8.85418781762e-12 * e + 1.256637062e-6 * m;
    return 0;
}

// ems_energy without allocation or synchronisation: the partial sums go to
// the device's file-scope scratch, the result to the caller's device memory
int ems_energy_noalloc(ems_engine* h, double* d_out)
{
    if (!h || !d_out) return fail(EMS_E_ARG, "null argument");
    if (h->device < 0 || h->device >= kMaxDevices || !g_scratch_ptr[h->device])
        return fail(EMS_E_STATE, "the device scratch of this engine's device is not set up");
    HIPCHK(hipSetDevice(h->device));
    double* part = g_scratch_ptr[h->device];
    hipLaunchKernelGGL(k_energy, dim3(kEnergyBlocks), dim3(kBlock), 0, h->stream, h->V, h->I, h->cs, h->N[0], h->N[1],
                       h->N[2], part);
    hipLaunchKernelGGL(k_energy_total, dim3(1), dim3(64), 0, h->stream, part, kEnergyBlocks, d_out);
    HIPCHK(hipGetLastError());
    return 0;
}

int ems_last_timing(ems_engine* h, double* total_ms, double* volt_ms, double* curr_ms)
{
    if (!h) return fail(EMS_E_ARG, "null handle");
    if (total_ms) *total_ms = h->last_ms;
    if (volt_ms) *volt_ms = h->last_volt_ms;
    if (curr_ms) *curr_ms = h->last_curr_ms;
    return 0;
}

void ems_destroy(ems_engine* h)
{
    if (!h) return;
    hipSetDevice(h->device);
    if (h->stream) hipStreamSynchronize(h->stream);
    for (auto& G : h->graphs) {
        if (G.exec) hipGraphExecDestroy(G.exec);
        if (G.graph) hipGraphDestroy(G.graph);
    }
    for (void* p : h->allocs) hipFree(p);
    if (h->ev0) hipEventDestroy(h->ev0);
    if (h->ev1) hipEventDestroy(h->ev1);
    for (auto& e : h->kev) hipEventDestroy(e);
    if (h->stream) hipStreamDestroy(h->stream);
    delete h;
}


// ---------------------------------------------------------------------------
// batched engine
// ---------------------------------------------------------------------------
}  // extern "C"

struct ems_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<ems_engine*> eng;
    int pm = kNoPml;
    bool cmp = false;
    std::vector<BatchInst> h_inst;
    BatchInst* d_inst = nullptr;
    // voltage sources in the voltage half step (every instance has its
    // lists); their tile masks hold for linear tiles only
    bool vsrc = false;
    std::vector<const unsigned*> vmask;
    // per-instance host lists, assembled into per-group device tables
    std::vector<std::vector<BatchExcGroup>> igrp[2];   // beg/end local to the instance's amp list
    std::vector<std::vector<float>> iamp[2];
    std::vector<std::vector<unsigned>> idel[2];
    std::vector<std::vector<BatchProbe>> iprobes;
    std::vector<size_t> ibytes;                         // device working set of each instance
    // groups of instances stepped together (all steps of a group run before
    // the next group starts, so a group that fits the last-level cache stays
    // resident); every group has its own work table and source / probe lists
    struct Group {
        std::vector<unsigned> inst;
        uint2* d_items = nullptr;
        unsigned nitems = 0, grid = 0;
        BatchExcGroup* d_grp[2] = {nullptr, nullptr};
        float* d_amp[2] = {nullptr, nullptr};
        unsigned* d_delay[2] = {nullptr, nullptr};
        unsigned G[2] = {0, 0};
        BatchProbe* d_probes = nullptr;
        unsigned nprobes = 0;
    };
    std::vector<Group> groups;
    std::vector<void*> gallocs;                         // group tables of the current configuration
    // each group's launches of graph_steps steps, captured once as a HIP graph
    // and replayed (0: direct launches, the default: replays measured no
    // faster on the batch, whose launches are long enough to hide the
    // launch overhead); dropped when the configuration changes
    unsigned graph_steps = 0;
    std::vector<hipGraphExec_t> gexec;
    // groups run on nstreams streams (group i on stream i % nstreams), so the
    // short source / probe launches and the last waves of one group's half
    // step overlap another group's work; stream 0 is `stream`
    unsigned nstreams = 1;
    std::vector<hipStream_t> xs;
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> joins;
    hipStream_t gstream(size_t gi) const { return gi % nstreams == 0 ? stream : xs[gi % nstreams - 1]; }
    unsigned tx = 0, grid_req = 0;
    size_t group_bytes = 0;
    std::vector<unsigned> zdepth;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0;
    int resident = 0;                                    // workgroups resident per CU (occupancy)
    int cus = 0;
    std::vector<void*> allocs;

    template <typename T>
    int alloc(T** p, size_t n)
    {
        void* q = nullptr;
        hipError_t e = hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T));
        if (e != hipSuccess) return fail(EMS_E_NOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
        allocs.push_back(q);
        *p = (T*)q;
        return 0;
    }
    template <typename T>
    int gupload(T** d, const T* src, size_t n)
    {
        void* q = nullptr;
        hipError_t e = hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T));
        if (e != hipSuccess) return fail(EMS_E_NOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
        gallocs.push_back(q);
        *d = (T*)q;
        if (n) HIPCHK(hipMemcpy(q, src, n * sizeof(T), hipMemcpyHostToDevice));
        return 0;
    }
};

namespace {

template <int PM, bool CC, bool CURR>
void batch_launch_t(ems_batch* b, const ems_batch::Group& g, hipStream_t st)
{
    hipLaunchKernelGGL((k_batch_half<PM, CC, CURR>), dim3(g.grid), dim3(kBlock), 0, st, b->d_inst, g.d_items,
                       g.nitems);
}

template <bool CURR>
void batch_launch(ems_batch* b, const ems_batch::Group& g, hipStream_t st)
{
    if (b->cmp) {
        if (b->pm == kPmlCompressed) batch_launch_t<kPmlCompressed, true, CURR>(b, g, st);
        else if (b->pm == kPmlPlain) batch_launch_t<kPmlPlain, true, CURR>(b, g, st);
        else batch_launch_t<kNoPml, true, CURR>(b, g, st);
    } else {
        if (b->pm == kPmlCompressed) batch_launch_t<kPmlCompressed, false, CURR>(b, g, st);
        else if (b->pm == kPmlPlain) batch_launch_t<kPmlPlain, false, CURR>(b, g, st);
        else batch_launch_t<kNoPml, false, CURR>(b, g, st);
    }
}

// Configuration: tile shape, z depth per instance, grid size, and the
// working-set budget of a group (0: all instances in one group).  Groups are
// packed first-fit by decreasing working set.
void batch_drop_graphs(ems_batch* b)
{
    for (auto& x : b->gexec)
        if (x) hipGraphExecDestroy(x);
    b->gexec.clear();
}

int batch_build(ems_batch* b, unsigned tx, const std::vector<unsigned>& zd, unsigned grid, size_t group_bytes)
{
    const size_t n = b->eng.size();
    HIPCHK(hipStreamSynchronize(b->stream));
    batch_drop_graphs(b);
    for (void* p : b->gallocs) hipFree(p);
    b->gallocs.clear();
    b->groups.clear();
    for (size_t i = 0; i < n; ++i) {
        BatchInst& B = b->h_inst[i];
        B.g.zdepth = zd[i];
        set_tiles(B.g, tx);
        if (b->vsrc) B.vsrc.tmask = tx == 0 ? b->vmask[i] : nullptr;
        if ((unsigned long long)B.g.qtiles * B.g.nzc > 0xFFFFFFFFull)
            return fail(EMS_E_ARG, "instance too large for the batch work table");
    }
    HIPCHK(hipMemcpy(b->d_inst, b->h_inst.data(), n * sizeof(BatchInst), hipMemcpyHostToDevice));
    // partition
    std::vector<std::vector<unsigned>> parts;
    if (group_bytes == 0) {
        parts.emplace_back();
        for (unsigned i = 0; i < n; ++i) parts[0].push_back(i);
    } else {
        std::vector<unsigned> order(n);
        for (unsigned i = 0; i < n; ++i) order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](unsigned a, unsigned c) { return b->ibytes[a] > b->ibytes[c]; });
        std::vector<size_t> load;
        for (unsigned i : order) {
            size_t k = 0;
            while (k < parts.size() && load[k] + b->ibytes[i] > group_bytes) ++k;
            if (k == parts.size()) {
                parts.emplace_back();
                load.push_back(0);
            }
            parts[k].push_back(i);
            load[k] += b->ibytes[i];
        }
        for (auto& p : parts) std::sort(p.begin(), p.end());
    }
    for (const auto& p : parts) {
        ems_batch::Group g;
        g.inst = p;
        std::vector<uint2> items;
        for (unsigned i : p) {
            const BatchInst& B = b->h_inst[i];
            const unsigned cnt = B.g.qtiles * B.g.nzc;
            for (unsigned t = 0; t < cnt; ++t) items.push_back(make_uint2(i, t));
        }
        int rc;
        if ((rc = b->gupload(&g.d_items, items.data(), items.size()))) return rc;
        g.nitems = (unsigned)items.size();
        g.grid = grid == 0 ? g.nitems : std::min(grid, g.nitems);
        for (int c = 0; c < 2; ++c) {
            std::vector<BatchExcGroup> grp;
            std::vector<float> amp;
            std::vector<unsigned> del;
            for (unsigned i : p) {
                for (BatchExcGroup q : b->igrp[c][i]) {
                    q.beg += (unsigned)amp.size();
                    q.end += (unsigned)amp.size();
                    grp.push_back(q);
                }
                amp.insert(amp.end(), b->iamp[c][i].begin(), b->iamp[c][i].end());
                del.insert(del.end(), b->idel[c][i].begin(), b->idel[c][i].end());
            }
            g.G[c] = (unsigned)grp.size();
            if (g.G[c] && ((rc = b->gupload(&g.d_grp[c], grp.data(), grp.size())) ||
                           (rc = b->gupload(&g.d_amp[c], amp.data(), amp.size())) ||
                           (rc = b->gupload(&g.d_delay[c], del.data(), del.size()))))
                return rc;
        }
        std::vector<BatchProbe> pr;
        for (unsigned i : p) pr.insert(pr.end(), b->iprobes[i].begin(), b->iprobes[i].end());
        g.nprobes = (unsigned)pr.size();
        if (g.nprobes && (rc = b->gupload(&g.d_probes, pr.data(), pr.size()))) return rc;
        b->groups.push_back(g);
    }
    b->tx = tx;
    b->zdepth = zd;
    b->grid_req = grid;
    b->group_bytes = group_bytes;
    return 0;
}

void batch_step(ems_batch* b, const ems_batch::Group& g, hipStream_t st)
{
    const unsigned T = 256;
    batch_launch<false>(b, g, st);
    if (g.G[0] && !b->vsrc)
        hipLaunchKernelGGL(k_batch_excite, dim3((g.G[0] + T - 1) / T), dim3(T), 0, st, b->d_inst, g.d_grp[0],
                           g.G[0], g.d_amp[0], g.d_delay[0], 0, 0);
    batch_launch<true>(b, g, st);
    if (g.G[1])
        hipLaunchKernelGGL(k_batch_excite, dim3((g.G[1] + T - 1) / T), dim3(T), 0, st, b->d_inst, g.d_grp[1],
                           g.G[1], g.d_amp[1], g.d_delay[1], 1, -1);
    if (g.nprobes)
        hipLaunchKernelGGL(k_batch_probe, dim3(g.nprobes), dim3(kBlock), 0, st, b->d_inst, g.d_probes);
}

// All kernel arguments of a step are the same at every step (the timestep
// is read on the device), so a chunk of steps replays as a graph.
int batch_run(ems_batch* b, unsigned nsteps)
{
    const unsigned S = b->graph_steps;
    if (S && b->gexec.size() != b->groups.size()) b->gexec.assign(b->groups.size(), nullptr);
    if (b->nstreams > 1) {
        // fork: the other streams start after what is queued on stream 0
        HIPCHK(hipEventRecord(b->fork, b->stream));
        for (auto st : b->xs) HIPCHK(hipStreamWaitEvent(st, b->fork, 0));
    }
    for (size_t gi = 0; gi < b->groups.size(); ++gi) {
        const auto& g = b->groups[gi];
        hipStream_t st = b->gstream(gi);
        unsigned s = 0;
        if (S && nsteps >= S) {
            if (!b->gexec[gi]) {
                hipGraph_t graph;
                HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
                for (unsigned k = 0; k < S; ++k) batch_step(b, g, st);
                HIPCHK(hipStreamEndCapture(st, &graph));
                hipError_t e = hipGraphInstantiate(&b->gexec[gi], graph, nullptr, nullptr, 0);
                hipGraphDestroy(graph);
                if (e != hipSuccess) {
                    b->gexec[gi] = nullptr;
                    return fail(EMS_E_HIP, std::string("graph instantiation failed: ") + hipGetErrorString(e));
                }
            }
            for (; s + S <= nsteps; s += S) HIPCHK(hipGraphLaunch(b->gexec[gi], st));
        }
        for (; s < nsteps; ++s) batch_step(b, g, st);
    }
    if (b->nstreams > 1)
        // join: stream 0 continues after every other stream's work
        for (size_t k = 0; k < b->xs.size(); ++k) {
            HIPCHK(hipEventRecord(b->joins[k], b->xs[k]));
            HIPCHK(hipStreamWaitEvent(b->stream, b->joins[k], 0));
        }
    return 0;
}

}  // namespace

extern "C" {

int ems_batch_create(ems_batch** out, ems_engine* const* engines, unsigned count)
{
    if (!out || !engines || count == 0) return fail(EMS_E_ARG, "null argument or empty batch");
    *out = nullptr;
    const ems_engine* e0 = engines[0];
    if (!e0) return fail(EMS_E_ARG, "null engine");
    for (unsigned i = 0; i < count; ++i) {
        const ems_engine* e = engines[i];
        if (!e) return fail(EMS_E_ARG, "null engine");
        if (!e->inited) return fail(EMS_E_STATE, "batch engines must be initialised (ems_init)");
        if (e->device != e0->device) return fail(EMS_E_ARG, "batch engines must share one device");
        if (e->cmp != e0->cmp) return fail(EMS_E_ARG, "batch engines must share one coefficient layout");
        const int pm = e->pmlV.n ? e->pml_mode : kNoPml, pm0 = e0->pmlV.n ? e0->pml_mode : kNoPml;
        if (pm != pm0) return fail(EMS_E_ARG, "batch engines must share one UPML layout (all with or all without)");
        if (!e->mur.empty() || !e->lor.empty() || !e->rlc.empty() || !e->dumps.empty() || e->has_ss)
            return fail(EMS_E_ARG, "batched instances support UPML, sources and probes only");
        if (e->fused) return fail(EMS_E_ARG, "batched instances run two half steps (ems_set_fused(h, 0))");
        if (e->numTS != e0->numTS) return fail(EMS_E_STATE, "batch engines must be at the same timestep");
        for (unsigned j = 0; j < i; ++j)
            if (engines[j] == e) return fail(EMS_E_ARG, "an engine appears twice in the batch");
    }
    HIPCHK(hipSetDevice(e0->device));
    ems_batch* b = new ems_batch();
    auto bail = [&](int r) {
        ems_batch_destroy(b);
        return r;
    };
    b->device = e0->device;
    b->eng.assign(engines, engines + count);
    b->cmp = e0->cmp;
    b->pm = e0->pmlV.n ? e0->pml_mode : kNoPml;
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(EMS_E_HIP, "stream creation failed"));
    hipEventCreate(&b->ev0);
    hipEventCreate(&b->ev1);
    int rc;
    for (int c = 0; c < 2; ++c) {
        b->igrp[c].resize(count);
        b->iamp[c].resize(count);
        b->idel[c].resize(count);
    }
    b->iprobes.resize(count);
    for (unsigned i = 0; i < count; ++i) {
        ems_engine* e = engines[i];
        HIPCHK(hipStreamSynchronize(e->stream));
        if ((rc = sync_probe_tables(e))) return bail(rc);
        BatchInst B{};
        B.g = e->g;
        B.V = e->V;
        B.I = e->I;
        B.kV = e->kV;
        B.kI = e->kI;
        B.pV = e->pmlV;
        B.pI = e->pmlI;
        B.numts = e->numts_dev;
        B.sig[0] = e->sig[0];
        B.sig[1] = e->sig[1];
        B.sigLen = e->sigLen;
        B.period = e->period;
        B.vsrc = e->quad ? ExcArgs{} : e->qsrc;
        b->h_inst.push_back(B);
        b->zdepth.push_back(e->g.zdepth);
        // device working set: fields, coefficients, UPML
        size_t bytes = 2 * 3 * e->cs * sizeof(float);
        bytes += e->cmp ? e->cs * sizeof(unsigned short) + 2 * 32 * (size_t)e->n_unique : 4 * 3 * e->cs * sizeof(float);
        if (e->pmlV.n) {
            const size_t np = e->pmlV.npos;
            bytes += 2 * 3 * np * sizeof(float);
            bytes += e->pml_mode == kPmlCompressed ? np * sizeof(unsigned) + 2 * 48 * (size_t)e->n_unique_pml
                                                   : 2 * 9 * np * sizeof(float);
        }
        b->ibytes.push_back(bytes);
        for (int c = 0; c < 2; ++c) {
            const auto& E = e->exc[c];
            for (unsigned g = 0; g < E.G; ++g) {
                BatchExcGroup q{};
                q.inst = i;
                q.lin = E.h_glin[g];
                q.beg = E.h_gbeg[g];
                q.end = E.h_gbeg[g + 1];
                b->igrp[c][i].push_back(q);
            }
            b->iamp[c][i] = E.h_amp;
            b->idel[c][i] = E.h_delay;
        }
        for (const auto& P : e->probes) {
            BatchProbe q{};
            q.P = P;
            q.inst = i;
            q.elin = e->d_elin;
            q.esign = e->d_esign;
            q.ediv = e->d_ediv;
            b->iprobes[i].push_back(q);
        }
    }
    b->vsrc = true;
    for (const auto& B : b->h_inst) b->vsrc = b->vsrc && B.vsrc.beg;
    for (auto& B : b->h_inst) {
        if (!b->vsrc) B.vsrc = ExcArgs{};
        b->vmask.push_back(B.vsrc.tmask);
    }
    if ((rc = b->alloc(&b->d_inst, count))) return bail(rc);
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, b->device));
    b->cus = prop.multiProcessorCount;
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_batch_half<kPmlCompressed, true, false>, kBlock, 0) !=
            hipSuccess ||
        occ <= 0)
        occ = 4;
    b->resident = occ;
    // default: each instance's own depth, linear tiles, one workgroup per item, one group
    if ((rc = batch_build(b, 0, b->zdepth, 0, 0))) return bail(rc);
    *out = b;
    return 0;
}

int ems_batch_set_config(ems_batch* b, unsigned tile_x, const unsigned* zdepth, unsigned grid)
{
    if (!b) return fail(EMS_E_ARG, "null batch");
    if (tile_x != 0 && (tile_x > kBlock || kBlock % tile_x || tile_x < 8))
        return fail(EMS_E_ARG, "tile_x must be 0 or a divisor of 256 from 8 to 256");
    std::vector<unsigned> zd(b->eng.size());
    for (size_t i = 0; i < zd.size(); ++i) {
        zd[i] = zdepth ? zdepth[i] : choose_zdepth(b->eng[i]->N);
        if (zd[i] < 1 || zd[i] > 4096) return fail(EMS_E_ARG, "z depth out of range");
    }
    return batch_build(b, tile_x, zd, grid, b->group_bytes);
}

int ems_batch_set_grouping(ems_batch* b, unsigned long long group_bytes)
{
    if (!b) return fail(EMS_E_ARG, "null batch");
    return batch_build(b, b->tx, b->zdepth, b->grid_req, (size_t)group_bytes);
}

int ems_batch_get_config(const ems_batch* b, unsigned* tile_x, unsigned* zdepth, unsigned* grid, unsigned* nitems)
{
    if (!b) return fail(EMS_E_ARG, "null batch");
    if (tile_x) *tile_x = b->tx;
    if (zdepth)
        for (size_t i = 0; i < b->zdepth.size(); ++i) zdepth[i] = b->zdepth[i];
    if (grid) *grid = b->grid_req;
    if (nitems) {
        unsigned t = 0;
        for (const auto& g : b->groups) t += g.nitems;
        *nitems = t;
    }
    return 0;
}

int ems_batch_get_grouping(const ems_batch* b, unsigned long long* group_bytes, unsigned* ngroups)
{
    if (!b) return fail(EMS_E_ARG, "null batch");
    if (group_bytes) *group_bytes = b->group_bytes;
    if (ngroups) *ngroups = (unsigned)b->groups.size();
    return 0;
}

int ems_batch_iterate(ems_batch* b, unsigned nsteps)
{
    if (!b) return fail(EMS_E_ARG, "null batch");
    TxRange tx("ems_batch_iterate");
    HIPCHK(hipSetDevice(b->device));
    for (auto* e : b->eng) HIPCHK(hipStreamSynchronize(e->stream));
    if (b->eng[0]->numTS == 0 && nsteps > 0)
        // Process() before the first step (openems.cpp:1424-1425)
        for (const auto& g : b->groups)
            if (g.nprobes)
                hipLaunchKernelGGL(k_batch_probe, dim3(g.nprobes), dim3(kBlock), 0, b->stream, b->d_inst, g.d_probes);
    HIPCHK(hipEventRecord(b->ev0, b->stream));
    int rc = batch_run(b, nsteps);
    if (rc) return rc;
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(b->ev1, b->stream));
    HIPCHK(hipEventSynchronize(b->ev1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    b->last_ms = ms;
    for (auto* e : b->eng) e->numTS += nsteps;
    return 0;
}

// Autotuner: time every candidate tile shape x z-depth policy x grid size x
// (stream count, group budget) on the batch itself for `steps` steps of all
// groups, restore the state the candidates advanced (fields, UPML fluxes,
// counters, probe sample counts), and keep the fastest.  Depth policies: each
// instance's own (choose_zdepth), 4 or 8 for all; grids: one workgroup per
// work item or the resident workgroups; plans: one stream with one group or
// budgets of 128-224 MiB, two streams with 80-112 MiB, three with 56-64 MiB
// (the last-level cache holds 256 MiB; the budget also sets how many work
// items a group's launches spread over the resident workgroups).
int ems_batch_autotune(ems_batch* b, unsigned steps, double* best_ms)
{
    if (!b) return fail(EMS_E_ARG, "null batch");
    if (steps < 2) steps = 2;
    HIPCHK(hipSetDevice(b->device));
    for (auto* e : b->eng) HIPCHK(hipStreamSynchronize(e->stream));
    struct Saved {
        float *V, *I, *fV, *fI;
        unsigned numts[2];
        std::vector<unsigned> counts;
    };
    std::vector<Saved> sv(b->eng.size());
    std::vector<void*> tmp;
    auto cleanup = [&]() {
        for (void* p : tmp) (void)hipFree(p);
    };
    auto dalloc = [&](float** p, size_t n) -> bool {
        void* q = nullptr;
        if (hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(float)) != hipSuccess) return false;
        tmp.push_back(q);
        *p = (float*)q;
        return true;
    };
    // every snapshot / restore copy is checked; the first failure is kept
    hipError_t cerr = hipSuccess;
    const char* cwhat = "";
    auto cp = [&](void* dst, const void* src, size_t bytes, hipMemcpyKind kind, const char* what) {
        const hipError_t e = hipMemcpy(dst, src, bytes, kind);
        if (e != hipSuccess && cerr == hipSuccess) {
            cerr = e;
            cwhat = what;
        }
    };
    auto copy_failed = [&](const char* phase) {
        return fail(EMS_E_HIP, std::string("ems_batch_autotune: ") + phase + " copy of " + cwhat + " failed: " +
                                   hipGetErrorString(cerr));
    };
    for (size_t i = 0; i < b->eng.size(); ++i) {
        ems_engine* e = b->eng[i];
        Saved& S = sv[i];
        S.fV = S.fI = nullptr;
        if (!dalloc(&S.V, 3 * e->cs) || !dalloc(&S.I, 3 * e->cs)) {
            cleanup();
            return fail(EMS_E_NOMEM, "no device memory for the autotuner's snapshot");
        }
        cp(S.V, e->V, 3 * e->cs * sizeof(float), hipMemcpyDeviceToDevice, "voltages");
        cp(S.I, e->I, 3 * e->cs * sizeof(float), hipMemcpyDeviceToDevice, "currents");
        if (e->pmlV.n) {
            if (!dalloc(&S.fV, 3 * e->pmlV.npos) || !dalloc(&S.fI, 3 * e->pmlI.npos)) {
                cleanup();
                return fail(EMS_E_NOMEM, "no device memory for the autotuner's snapshot");
            }
            cp(S.fV, e->pmlV.flux_in, 3 * e->pmlV.npos * sizeof(float), hipMemcpyDeviceToDevice, "voltage flux");
            cp(S.fI, e->pmlI.flux_in, 3 * e->pmlI.npos * sizeof(float), hipMemcpyDeviceToDevice, "current flux");
        }
        cp(S.numts, e->numts_dev, 2 * sizeof(unsigned), hipMemcpyDeviceToHost, "step counter");
        for (const auto& P : e->probes) {
            unsigned c = 0;
            cp(&c, P.count, sizeof(unsigned), hipMemcpyDeviceToHost, "probe count");
            S.counts.push_back(c);
        }
    }
    if (cerr != hipSuccess) {
        cleanup();
        return copy_failed("snapshot");
    }
    {
        const hipError_t e = hipDeviceSynchronize();
        if (e != hipSuccess) {
            cleanup();
            return fail(EMS_E_HIP, std::string("ems_batch_autotune: ") + hipGetErrorString(e));
        }
    }
    // puts every instance back to the snapshot; returns 0 or an EMS_E_HIP
    // status (the instances are then in an undefined state)
    auto restore = [&]() -> int {
        for (size_t i = 0; i < b->eng.size(); ++i) {
            ems_engine* e = b->eng[i];
            Saved& S = sv[i];
            cp(e->V, S.V, 3 * e->cs * sizeof(float), hipMemcpyDeviceToDevice, "voltages");
            cp(e->I, S.I, 3 * e->cs * sizeof(float), hipMemcpyDeviceToDevice, "currents");
            if (e->pmlV.n) {
                cp(const_cast<float*>(e->pmlV.flux_in), S.fV, 3 * e->pmlV.npos * sizeof(float),
                   hipMemcpyDeviceToDevice, "voltage flux");
                cp(const_cast<float*>(e->pmlI.flux_in), S.fI, 3 * e->pmlI.npos * sizeof(float),
                   hipMemcpyDeviceToDevice, "current flux");
            }
            cp(e->numts_dev, S.numts, 2 * sizeof(unsigned), hipMemcpyHostToDevice, "step counter");
            for (size_t k = 0; k < e->probes.size(); ++k)
                cp(e->probes[k].count, &S.counts[k], sizeof(unsigned), hipMemcpyHostToDevice, "probe count");
        }
        // device-to-device copies are not ordered with the batch's stream
        const hipError_t se = hipDeviceSynchronize();
        if (se != hipSuccess && cerr == hipSuccess) {
            cerr = se;
            cwhat = "(synchronisation)";
        }
        return cerr == hipSuccess ? 0 : copy_failed("restore");
    };
    const unsigned txs[2] = {0, 64};
    const unsigned depths[3] = {0, 4, 8};   // 0: each instance's own choice
    const unsigned grids[2] = {0, (unsigned)(b->resident * b->cus)};
    const size_t mib = size_t(1) << 20;
    // (streams, group budget): one stream with budgets up to the cache size,
    // several streams with a budget near the cache size over the stream count
    const std::pair<unsigned, size_t> plans[] = {{1, 0},        {1, 128 * mib}, {1, 160 * mib}, {1, 192 * mib},
                                                 {1, 224 * mib}, {2, 80 * mib},  {2, 96 * mib},  {2, 112 * mib},
                                                 {3, 56 * mib},  {3, 64 * mib}};
    const unsigned streams0 = b->nstreams;
    double best = 1e30;
    unsigned btx = 0, bdep = 0, bgrid = 0, bns = streams0;
    size_t bbud = 0;
    for (const auto& plan : plans)
        for (unsigned tx : txs)
            for (unsigned dep : depths)
                for (unsigned grid : grids) {
                    std::vector<unsigned> zd(b->eng.size());
                    for (size_t i = 0; i < zd.size(); ++i) zd[i] = dep ? dep : choose_zdepth(b->eng[i]->N);
                    int rc = batch_build(b, tx, zd, grid, plan.second);
                    if (!rc) rc = ems_batch_set_streams(b, plan.first);
                    if (rc) {
                        const int rr = restore();
                        cleanup();
                        return rr ? rr : rc;
                    }
                    // one warm-up step of every group, then steps-1 timed steps of all groups
                    if ((rc = batch_run(b, 1))) {
                        const int rr = restore();
                        cleanup();
                        return rr ? rr : rc;
                    }
                    hipError_t te = hipEventRecord(b->ev0, b->stream);
                    rc = batch_run(b, steps - 1);
                    if (te == hipSuccess) te = hipEventRecord(b->ev1, b->stream);
                    if (te == hipSuccess) te = hipEventSynchronize(b->ev1);
                    float ms = 0;
                    if (te == hipSuccess) te = hipEventElapsedTime(&ms, b->ev0, b->ev1);
                    if (!rc && te != hipSuccess)
                        rc = fail(EMS_E_HIP, std::string("ems_batch_autotune: timing: ") + hipGetErrorString(te));
                    if (rc) {
                        const int rr = restore();
                        cleanup();
                        return rr ? rr : rc;
                    }
                    const double tot = ms / (steps - 1);
                    if (tot < best) {
                        best = tot;
                        btx = tx;
                        bdep = dep;
                        bgrid = grid;
                        bbud = plan.second;
                        bns = plan.first;
                    }
                    if ((rc = restore())) {
                        cleanup();
                        return rc;
                    }
                }
    std::vector<unsigned> zd(b->eng.size());
    for (size_t i = 0; i < zd.size(); ++i) zd[i] = bdep ? bdep : choose_zdepth(b->eng[i]->N);
    int rc = batch_build(b, btx, zd, bgrid, bbud);
    if (!rc) rc = ems_batch_set_streams(b, bns);
    const hipError_t se = hipStreamSynchronize(b->stream);
    cleanup();
    if (!rc && se != hipSuccess)
        rc = fail(EMS_E_HIP, std::string("ems_batch_autotune: ") + hipGetErrorString(se));
    if (best_ms) *best_ms = best;
    return rc;
}

int ems_batch_last_ms(const ems_batch* b, double* ms)
{
    if (!b || !ms) return fail(EMS_E_ARG, "null argument");
    *ms = b->last_ms;
    return 0;
}

int ems_batch_set_graph(ems_batch* b, unsigned steps)
{
    if (!b) return fail(EMS_E_ARG, "null batch");
    if (steps > 4096) return fail(EMS_E_ARG, "graph chunk longer than 4096 steps");
    HIPCHK(hipSetDevice(b->device));
    HIPCHK(hipStreamSynchronize(b->stream));
    batch_drop_graphs(b);
    b->graph_steps = steps;
    return 0;
}

int ems_batch_set_streams(ems_batch* b, unsigned n)
{
    if (!b) return fail(EMS_E_ARG, "null batch");
    if (n < 1 || n > 4) return fail(EMS_E_ARG, "stream count must be 1..4");
    HIPCHK(hipSetDevice(b->device));
    HIPCHK(hipStreamSynchronize(b->stream));
    for (auto st : b->xs) HIPCHK(hipStreamSynchronize(st));
    batch_drop_graphs(b);       // captured on the streams of the old assignment
    if (!b->fork) HIPCHK(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
    while (b->xs.size() < n - 1) {
        hipStream_t st;
        HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        b->xs.push_back(st);
        hipEvent_t ev;
        HIPCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        b->joins.push_back(ev);
    }
    while (b->xs.size() > n - 1) {
        hipStreamDestroy(b->xs.back());
        hipEventDestroy(b->joins.back());
        b->xs.pop_back();
        b->joins.pop_back();
    }
    b->nstreams = n;
    return 0;
}

int ems_batch_get_streams(const ems_batch* b, unsigned* n)
{
    if (!b || !n) return fail(EMS_E_ARG, "null argument");
    *n = b->nstreams;
    return 0;
}

int ems_batch_get_graph(const ems_batch* b, unsigned* steps)
{
    if (!b || !steps) return fail(EMS_E_ARG, "null argument");
    *steps = b->graph_steps;
    return 0;
}

void ems_batch_destroy(ems_batch* b)
{
    if (!b) return;
    hipSetDevice(b->device);
    if (b->stream) hipStreamSynchronize(b->stream);
    for (auto st : b->xs) hipStreamSynchronize(st);
    batch_drop_graphs(b);
    for (auto st : b->xs) hipStreamDestroy(st);
    for (auto ev : b->joins) hipEventDestroy(ev);
    if (b->fork) hipEventDestroy(b->fork);
    for (void* p : b->gallocs) hipFree(p);
    for (void* p : b->allocs) hipFree(p);
    if (b->ev0) hipEventDestroy(b->ev0);
    if (b->ev1) hipEventDestroy(b->ev1);
    if (b->stream) hipStreamDestroy(b->stream);
    delete b;
}


// ---------------------------------------------------------------------------
// adjoint gradients
// ---------------------------------------------------------------------------
}  // extern "C"

namespace {

// one saved state: fields, UPML fluxes, timestep
struct AdjCkpt {
    float *V = nullptr, *I = nullptr, *FV = nullptr, *FI = nullptr;
    unsigned ts = 0;
};

}  // namespace

struct ems_adjoint {
    ems_engine* h = nullptr;
    AdjProbe* d_probes = nullptr;
    unsigned np = 0;
    double freq = 0, dT = 0;
    AdjBox box{};
    unsigned K = 256;
    unsigned n0 = 0, nsteps = 0;
    double* d_F = nullptr;
    std::vector<double> F;
    std::vector<AdjCkpt> ck;   // state at n0 + s*K, s = 0..S-1, then the final state
    float *lV = nullptr, *lI = nullptr, *A = nullptr, *B = nullptr, *lFV = nullptr, *lFI = nullptr;
    float *hc = nullptr, *hv = nullptr;   // [K][3][box.n]
    double *gvi = nullptr, *gvv = nullptr;
    bool forward_done = false;
    std::vector<void*> allocs;
    template <typename T>
    int alloc(T** p, size_t n)
    {
        void* q = nullptr;
        hipError_t e = hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T));
        if (e != hipSuccess) return fail(EMS_E_NOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
        allocs.push_back(q);
        *p = (T*)q;
        return 0;
    }
};

namespace {

int adj_save(ems_adjoint* a, AdjCkpt& c)
{
    ems_engine* h = a->h;
    const size_t fb = 3 * h->cs * sizeof(float);
    int rc;
    if (!c.V) {
        if ((rc = a->alloc(&c.V, 3 * h->cs)) || (rc = a->alloc(&c.I, 3 * h->cs))) return rc;
        if (h->pmlV.n && ((rc = a->alloc(&c.FV, 3 * h->pmlV.npos)) || (rc = a->alloc(&c.FI, 3 * h->pmlI.npos))))
            return rc;
    }
    HIPCHK(hipMemcpyAsync(c.V, h->V, fb, hipMemcpyDeviceToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(c.I, h->I, fb, hipMemcpyDeviceToDevice, h->stream));
    if (h->pmlV.n) {
        HIPCHK(hipMemcpyAsync(c.FV, h->pmlV.flux_in, 3 * h->pmlV.npos * sizeof(float), hipMemcpyDeviceToDevice,
                              h->stream));
        HIPCHK(hipMemcpyAsync(c.FI, h->pmlI.flux_in, 3 * h->pmlI.npos * sizeof(float), hipMemcpyDeviceToDevice,
                              h->stream));
    }
    c.ts = h->numTS;
    return 0;
}

int adj_restore(ems_adjoint* a, const AdjCkpt& c)
{
    ems_engine* h = a->h;
    const size_t fb = 3 * h->cs * sizeof(float);
    HIPCHK(hipMemcpyAsync(h->V, c.V, fb, hipMemcpyDeviceToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(h->I, c.I, fb, hipMemcpyDeviceToDevice, h->stream));
    if (h->pmlV.n) {
        HIPCHK(hipMemcpyAsync(const_cast<float*>(h->pmlV.flux_in), c.FV, 3 * h->pmlV.npos * sizeof(float),
                              hipMemcpyDeviceToDevice, h->stream));
        HIPCHK(hipMemcpyAsync(const_cast<float*>(h->pmlI.flux_in), c.FI, 3 * h->pmlI.npos * sizeof(float),
                              hipMemcpyDeviceToDevice, h->stream));
    }
    const unsigned ts[2] = {c.ts, 0u};
    HIPCHK(hipMemcpyAsync(h->numts_dev, ts, sizeof(ts), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    h->numTS = c.ts;
    return 0;
}

template <int PM, bool CC>
void adj_kernels_t(ems_adjoint* a, const float* hc, const float* hv)
{
    ems_engine* h = a->h;
    const unsigned nblk = h->g.qtiles * h->g.nzc;
    PmlArgs pI = h->pmlI, pV = h->pmlV;
    hipLaunchKernelGGL((k_adj_curr<PM, CC>), dim3(nblk), dim3(kBlock), 0, h->stream, h->g, a->lI, a->B, a->A, h->kI,
                       pI, a->lFI);
    hipLaunchKernelGGL((k_adj_volt<PM, CC>), dim3(nblk), dim3(kBlock), 0, h->stream, h->g, a->lV, a->A, a->B, h->kV,
                       pV, a->lFV, a->box, hc, hv, a->gvi, a->gvv);
}

void adj_kernels(ems_adjoint* a, const float* hc, const float* hv)
{
    ems_engine* h = a->h;
    const int pm = h->pmlV.n ? h->pml_mode : kNoPml;
    if (h->cmp) {
        if (pm == kPmlCompressed) adj_kernels_t<kPmlCompressed, true>(a, hc, hv);
        else if (pm == kPmlPlain) adj_kernels_t<kPmlPlain, true>(a, hc, hv);
        else adj_kernels_t<kNoPml, true>(a, hc, hv);
    } else {
        if (pm == kPmlCompressed) adj_kernels_t<kPmlCompressed, false>(a, hc, hv);
        else if (pm == kPmlPlain) adj_kernels_t<kPmlPlain, false>(a, hc, hv);
        else adj_kernels_t<kNoPml, false>(a, hc, hv);
    }
}

}  // namespace

extern "C" {

int ems_adj_create(ems_adjoint** out, ems_engine* h, unsigned nprobe, const unsigned* comp, const unsigned* pos,
                   const double* divisor, double freq, double dT, const unsigned box_lo[3], const unsigned box_hi[3],
                   unsigned checkpoint_every)
{
    if (!out || !h || !comp || !pos || !divisor || !box_lo || !box_hi) return fail(EMS_E_ARG, "null argument");
    *out = nullptr;
    if (!h->inited) return fail(EMS_E_STATE, "initialise the engine (ems_init) first");
    if (h->fused) return fail(EMS_E_ARG, "the adjoint runs the two-half-step form (ems_set_fused(h, 0))");
    if (!h->mur.empty() || !h->lor.empty() || !h->rlc.empty())
        return fail(EMS_E_ARG, "the adjoint supports UPML and sources only (no Mur, Lorentz, lumped RLC)");
    if (nprobe == 0 || !(dT > 0)) return fail(EMS_E_ARG, "no objective probes or bad timestep");
    if (checkpoint_every == 0) return fail(EMS_E_ARG, "checkpoint interval must be positive");
    AdjBox box{};
    box.x0 = box_lo[0];
    box.y0 = box_lo[1];
    box.z0 = box_lo[2];
    for (int n = 0; n < 3; ++n)
        if (box_hi[n] < box_lo[n] || box_hi[n] >= h->N[n]) return fail(EMS_E_ARG, "sensitivity box outside the mesh");
    box.nx = box_hi[0] - box_lo[0] + 1;
    box.ny = box_hi[1] - box_lo[1] + 1;
    box.nz = box_hi[2] - box_lo[2] + 1;
    box.n = (unsigned long long)box.nx * box.ny * box.nz;
    // the box must stay clear of the UPML (its coefficients are not the ones differentiated)
    for (int s = 0; s < h->pmlV.n; ++s) {
        const SlabTab& S = h->pmlV.tab[s];
        if ((int)box_lo[0] <= S.x1 && (int)box_hi[0] >= S.x0 && (int)box_lo[1] <= S.y1 && (int)box_hi[1] >= S.y0 &&
            (int)box_lo[2] <= S.z1 && (int)box_hi[2] >= S.z0)
            return fail(EMS_E_ARG, "the sensitivity box overlaps the UPML");
    }
    std::vector<AdjProbe> pr(nprobe);
    for (unsigned p = 0; p < nprobe; ++p) {
        const unsigned x = pos[p], y = pos[nprobe + p], z = pos[2 * nprobe + p];
        if (x >= h->N[0] || y >= h->N[1] || z >= h->N[2] || comp[p] > 2 || !(divisor[p] != 0.0))
            return fail(EMS_E_ARG, "objective probe outside the mesh or zero divisor");
        pr[p].lin = comp[p] * h->cs + h->lin(x, y, z);
        pr[p].div = divisor[p];
    }
    HIPCHK(hipSetDevice(h->device));
    ems_adjoint* a = new ems_adjoint();
    auto bail = [&](int r) {
        ems_adj_destroy(a);
        return r;
    };
    a->h = h;
    a->np = nprobe;
    a->freq = freq;
    a->dT = dT;
    a->box = box;
    a->K = checkpoint_every;
    int rc;
    const size_t c3 = 3 * h->cs;
    const size_t f3V = h->pmlV.n ? 3 * h->pmlV.npos : 1, f3I = h->pmlI.n ? 3 * h->pmlI.npos : 1;
    if ((rc = a->alloc(&a->d_probes, nprobe)) || (rc = a->alloc(&a->d_F, 2 * nprobe)) || (rc = a->alloc(&a->lV, c3)) ||
        (rc = a->alloc(&a->lI, c3)) || (rc = a->alloc(&a->A, c3)) || (rc = a->alloc(&a->B, c3)) ||
        (rc = a->alloc(&a->lFV, f3V)) || (rc = a->alloc(&a->lFI, f3I)) ||
        (rc = a->alloc(&a->hc, (size_t)a->K * 3 * box.n)) || (rc = a->alloc(&a->hv, (size_t)a->K * 3 * box.n)) ||
        (rc = a->alloc(&a->gvi, 3 * box.n)) || (rc = a->alloc(&a->gvv, 3 * box.n)))
        return bail(rc);
    if (hipMemcpy(a->d_probes, pr.data(), nprobe * sizeof(AdjProbe), hipMemcpyHostToDevice) != hipSuccess)
        return bail(fail(EMS_E_HIP, "upload of the objective probes failed"));
    *out = a;
    return 0;
}

int ems_adj_forward(ems_adjoint* a, unsigned nsteps, double* J)
{
    TxRange tx("ems_adj_forward");
    if (!a || nsteps == 0) return fail(EMS_E_ARG, "null adjoint or no steps");
    ems_engine* h = a->h;
    HIPCHK(hipSetDevice(h->device));
    int rc;
    if ((rc = sync_probe_tables(h))) return rc;
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipMemsetAsync(a->d_F, 0, 2 * a->np * sizeof(double), h->stream));
    a->n0 = h->numTS;
    a->nsteps = nsteps;
    const unsigned S = (nsteps + a->K - 1) / a->K;
    if (a->ck.size() < (size_t)S + 1) a->ck.resize(S + 1);
    const unsigned T = 256;
    for (unsigned s = 0; s < S; ++s) {
        if ((rc = adj_save(a, a->ck[s]))) return rc;
        const unsigned len = std::min(a->K, nsteps - s * a->K);
        for (unsigned j = 0; j < len; ++j) {
            rc = launch_step(h, nullptr);
            ++h->numTS;
            if (rc) return rc;
            hipLaunchKernelGGL(k_adj_dft, dim3((a->np + T - 1) / T), dim3(T), 0, h->stream, h->V, a->d_probes, a->np,
                               a->freq, a->dT, h->numts_dev, a->d_F);
        }
    }
    if ((rc = adj_save(a, a->ck[S]))) return rc;
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->stream));
    a->F.assign(2 * a->np, 0.0);
    HIPCHK(hipMemcpy(a->F.data(), a->d_F, a->F.size() * sizeof(double), hipMemcpyDeviceToHost));
    double j = 0;
    for (unsigned p = 0; p < a->np; ++p) j += a->F[2 * p] * a->F[2 * p] + a->F[2 * p + 1] * a->F[2 * p + 1];
    if (J) *J = j;
    a->forward_done = true;
    return 0;
}

int ems_adj_backward(ems_adjoint* a, double* dJ_dvi, double* dJ_dvv)
{
    TxRange tx("ems_adj_backward");
    if (!a) return fail(EMS_E_ARG, "null adjoint");
    if (!a->forward_done) return fail(EMS_E_STATE, "run ems_adj_forward first");
    ems_engine* h = a->h;
    HIPCHK(hipSetDevice(h->device));
    const size_t c3 = 3 * h->cs * sizeof(float);
    HIPCHK(hipMemsetAsync(a->lV, 0, c3, h->stream));
    HIPCHK(hipMemsetAsync(a->lI, 0, c3, h->stream));
    HIPCHK(hipMemsetAsync(a->A, 0, c3, h->stream));
    HIPCHK(hipMemsetAsync(a->B, 0, c3, h->stream));
    if (h->pmlV.n) {
        HIPCHK(hipMemsetAsync(a->lFV, 0, 3 * h->pmlV.npos * sizeof(float), h->stream));
        HIPCHK(hipMemsetAsync(a->lFI, 0, 3 * h->pmlI.npos * sizeof(float), h->stream));
    }
    HIPCHK(hipMemsetAsync(a->gvi, 0, 3 * a->box.n * sizeof(double), h->stream));
    HIPCHK(hipMemsetAsync(a->gvv, 0, 3 * a->box.n * sizeof(double), h->stream));
    const unsigned S = (a->nsteps + a->K - 1) / a->K;
    const unsigned T = 256;
    const unsigned rblk = (unsigned)((a->box.n + T - 1) / T);
    int rc;
    for (int s = (int)S - 1; s >= 0; --s) {
        if ((rc = adj_restore(a, a->ck[s]))) return rc;
        const unsigned len = std::min(a->K, a->nsteps - s * a->K);
        // replay the segment, recording the inputs of each voltage update
        for (unsigned j = 0; j < len; ++j) {
            hipLaunchKernelGGL(k_adj_record, dim3(rblk), dim3(T), 0, h->stream, h->g, h->V, h->I, a->box,
                               a->hc + (size_t)j * 3 * a->box.n, a->hv + (size_t)j * 3 * a->box.n);
            rc = launch_step(h, nullptr);
            ++h->numTS;
            if (rc) return rc;
        }
        // reverse steps of the segment
        for (int j = (int)len - 1; j >= 0; --j) {
            const unsigned ts = a->n0 + (unsigned)s * a->K + (unsigned)j + 1;
            hipLaunchKernelGGL(k_adj_src, dim3((a->np + T - 1) / T), dim3(T), 0, h->stream, a->lV, a->d_probes,
                               a->np, a->freq, a->dT, ts, a->d_F);
            adj_kernels(a, a->hc + (size_t)j * 3 * a->box.n, a->hv + (size_t)j * 3 * a->box.n);
        }
    }
    HIPCHK(hipGetLastError());
    // leave the engine at the end of the forward run
    if ((rc = adj_restore(a, a->ck[S]))) return rc;
    if (dJ_dvi) HIPCHK(hipMemcpy(dJ_dvi, a->gvi, 3 * a->box.n * sizeof(double), hipMemcpyDeviceToHost));
    if (dJ_dvv) HIPCHK(hipMemcpy(dJ_dvv, a->gvv, 3 * a->box.n * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

void ems_adj_destroy(ems_adjoint* a)
{
    if (!a) return;
    if (a->h) {
        hipSetDevice(a->h->device);
        if (a->h->stream) hipStreamSynchronize(a->h->stream);
    }
    for (void* p : a->allocs) hipFree(p);
    delete a;
}

}  // extern "C"
