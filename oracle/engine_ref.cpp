// Oracle: CPU restatement of the openEMS EC-FDTD engines and engine extensions.
//
// TEST INFRASTRUCTURE ONLY.  The product (openems_amd) never links, loads or
// calls this file; tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg use it as the checker and as the restated-reference CPU baseline.
//
// Build: see oracle/Makefile (-O3 -msse2 -mno-fma -ffp-contract=off, so every
// float operation is a separately rounded IEEE op, as in the reference's
// x86-64 SSE2 build).
//
// Restated reference code paths:
//   basic engine        FDTD/engine.cpp:110-222 (UpdateVoltages/Currents), :224-286 (hooks, IterateTS)
//   extension ordering  FDTD/engine.cpp:87-98, FDTD/extensions/engine_extension.h:21-29
//   SSE engine          FDTD/engine_sse.cpp:79-274, lane layout FDTD/engine_sse.h:38-42,
//                       I-J-K-N storage tools/arraylib/array_e.h:53-61
//   SSE compressed      FDTD/operator_sse_compressed.cpp:114-225, FDTD/engine_sse_compressed.cpp:51-311
//   multithreaded       FDTD/engine_multithread.cpp:142-232,310-402, FDTD/operator_multithread.cpp:93-114,
//                       tools/useful.cpp:45-75 (AssignJobs2Threads)
//   excitation          FDTD/extensions/engine_ext_excitation.cpp:33-99
//   UPML                FDTD/extensions/engine_ext_upml.cpp:52-234
//   Mur ABC             FDTD/extensions/engine_ext_mur_abc.cpp:28-178
//   lumped RLC          FDTD/extensions/engine_ext_lumpedRLC.cpp:83-147
//   Lorentz/Drude ADE   FDTD/extensions/engine_ext_lorentzmaterial.cpp:79-173,
//                       FDTD/extensions/engine_ext_dispersive.cpp:76-132
//   energy              FDTD/engine_interface_fdtd.cpp:302-347
#include <algorithm>
#include <atomic>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <map>
#include <mutex>
#include <thread>
#include <vector>
#include <xmmintrin.h>
#include <pmmintrin.h>

typedef float v4sf __attribute__((vector_size(16)));
union f4vector {
    v4sf v;
    float f[4];
};

static std::vector<unsigned> assign_jobs(unsigned jobs, unsigned nthreads)
{
    // AssignJobs2Threads (tools/useful.cpp:45-75), RemoveEmpty=false
    std::vector<unsigned> jpt;
    unsigned per = jobs / nthreads;
    for (unsigned n = 0; n < nthreads; ++n) {
        jpt.push_back(per);
        jobs -= per;
    }
    for (unsigned n = 0; n < nthreads; ++n)
        if (jobs > 0) {
            ++jpt[n];
            --jobs;
        }
    return jpt;
}

struct Barrier {
    // boost::barrier equivalent (mutex + condition variable, generation count)
    explicit Barrier(unsigned n) : count(n), waiting(0), gen(0) {}
    void wait()
    {
        std::unique_lock<std::mutex> lk(m);
        unsigned g = gen;
        if (++waiting == count) {
            waiting = 0;
            ++gen;
            cv.notify_all();
            return;
        }
        cv.wait(lk, [&] { return g != gen; });
    }
    std::mutex m;
    std::condition_variable cv;
    unsigned count, waiting, gen;
};

// ---------------------------------------------------------------------------
// engine storage abstraction: every engine exposes GetVolt/SetVolt/GetCurr/SetCurr
// with the reference's (n, x, y, z) addressing, so the extensions below are
// written once (the reference does the same with ENG_DISPATCH,
// FDTD/extensions/engine_extension_dispatcher.h:58-88).
// ---------------------------------------------------------------------------
struct EngineBase {
    unsigned N[3];
    unsigned numTS = 0;
    virtual ~EngineBase() {}
    virtual double getV(int n, unsigned x, unsigned y, unsigned z) const = 0;
    virtual double getI(int n, unsigned x, unsigned y, unsigned z) const = 0;
    virtual void iterate(unsigned steps) = 0;
};

template <typename T>
struct Ext {
    double prio = 0;
    virtual ~Ext() {}
    // threadID-aware hooks; the default forwards to thread 0 only
    // (FDTD/extensions/engine_extension.cpp:50-90)
    virtual bool mt() const { return false; }
    virtual void setThreads(unsigned) {}
};

// --- generic field accessor bound to an engine type --------------------------
template <typename T, typename Eng>
struct ExtT : Ext<T> {
    virtual void preV(Eng&, unsigned) {}
    virtual void postV(Eng&, unsigned) {}
    virtual void applyV(Eng&, unsigned) {}
    virtual void preI(Eng&, unsigned) {}
    virtual void postI(Eng&, unsigned) {}
    virtual void applyI(Eng&, unsigned) {}
};

// --- excitation ---------------------------------------------------------------
template <typename T, typename Eng>
struct ExcExt : ExtT<T, Eng> {
    std::vector<unsigned> vidx[3], cidx[3], vdelay, cdelay;
    std::vector<unsigned short> vdir, cdir;
    std::vector<float> vamp, camp;
    std::vector<float> sigV, sigC;
    unsigned length = 0;
    int period_steps = 0;  // int(period/dT) or 0
    ExcExt() { this->prio = -1000; }
    void apply(Eng& e, bool curr)
    {
        // Engine_Ext_Excitation::Apply2Voltages/Apply2Current
        int numTS = (int)e.numTS;
        int p = numTS + 1;
        if (period_steps > 0) p = period_steps;
        const auto& idx = curr ? cidx : vidx;
        const auto& delay = curr ? cdelay : vdelay;
        const auto& dir = curr ? cdir : vdir;
        const auto& amp = curr ? camp : vamp;
        const auto& sig = curr ? sigC : sigV;
        for (size_t n = 0; n < dir.size(); ++n) {
            int exc_pos = numTS - (int)delay[n];
            exc_pos *= (exc_pos > 0);
            exc_pos %= p;
            exc_pos *= (exc_pos < (int)length);
            unsigned ny = dir[n];
            T v = curr ? e.GetCurr(ny, idx[0][n], idx[1][n], idx[2][n]) : e.GetVolt(ny, idx[0][n], idx[1][n], idx[2][n]);
            T add = (T)(amp[n] * sig[exc_pos]);
            if (curr)
                e.SetCurr(ny, idx[0][n], idx[1][n], idx[2][n], v + add);
            else
                e.SetVolt(ny, idx[0][n], idx[1][n], idx[2][n], v + add);
        }
    }
    void applyV(Eng& e, unsigned tid) override
    {
        if (tid == 0) apply(e, false);
    }
    void applyI(Eng& e, unsigned tid) override
    {
        if (tid == 0) apply(e, true);
    }
};

// --- UPML -----------------------------------------------------------------------
template <typename T, typename Eng>
struct UpmlExt : ExtT<T, Eng> {
    unsigned start[3], num[3];
    std::vector<T> vv, vvfo, vvfn, ii, iifo, iifn, vflux, cflux;
    std::vector<unsigned> numX, startX;
    UpmlExt() { this->prio = 1e6; }
    bool mt() const override { return true; }
    void setThreads(unsigned n) override
    {
        numX = assign_jobs(num[0], n);
        startX.assign(n, 0);
        for (size_t k = 1; k < numX.size(); ++k) startX[k] = startX[k - 1] + numX[k - 1];
    }
    size_t li(int n, unsigned x, unsigned y, unsigned z) const
    {
        return (((size_t)n * num[0] + x) * num[1] + y) * num[2] + z;
    }
    template <bool CURR, bool PRE>
    void hook(Eng& e, unsigned tid)
    {
        if (tid >= numX.size()) return;
        std::vector<T>& flux = CURR ? cflux : vflux;
        const std::vector<T>& c0 = CURR ? ii : vv;
        const std::vector<T>& cfo = CURR ? iifo : vvfo;
        const std::vector<T>& cfn = CURR ? iifn : vvfn;
        for (unsigned lx = startX[tid]; lx < startX[tid] + numX[tid]; ++lx)
            for (unsigned ly = 0; ly < num[1]; ++ly)
                for (unsigned lz = 0; lz < num[2]; ++lz) {
                    unsigned x = lx + start[0], y = ly + start[1], z = lz + start[2];
                    for (int n = 0; n < 3; ++n) {
                        size_t i = li(n, lx, ly, lz);
                        T f = CURR ? e.GetCurr(n, x, y, z) : e.GetVolt(n, x, y, z);
                        if (PRE) {
                            // f_help = vv*V - vvfo*flux; V = flux; flux = f_help
                            T help = c0[i] * f - cfo[i] * flux[i];
                            if (CURR) e.SetCurr(n, x, y, z, flux[i]); else e.SetVolt(n, x, y, z, flux[i]);
                            flux[i] = help;
                        } else {
                            // f_help = flux; flux = V; V = f_help + vvfn*flux
                            T help = flux[i];
                            flux[i] = f;
                            T nv = help + cfn[i] * flux[i];
                            if (CURR) e.SetCurr(n, x, y, z, nv); else e.SetVolt(n, x, y, z, nv);
                        }
                    }
                }
    }
    void preV(Eng& e, unsigned t) override { hook<false, true>(e, t); }
    void postV(Eng& e, unsigned t) override { hook<false, false>(e, t); }
    void preI(Eng& e, unsigned t) override { hook<true, true>(e, t); }
    void postI(Eng& e, unsigned t) override { hook<true, false>(e, t); }
};

// --- Mur ABC ----------------------------------------------------------------------
template <typename T, typename Eng>
struct MurExt : ExtT<T, Eng> {
    int ny, nyP, nyPP;
    unsigned line, line_shift, nl[2], start_ts;
    std::vector<T> cP, cPP, mP, mPP;
    std::vector<unsigned> numX, startX;
    MurExt() { this->prio = 0; }
    bool mt() const override { return true; }
    void setThreads(unsigned n) override
    {
        numX = assign_jobs(nl[0], n);
        startX.assign(n, 0);
        for (size_t k = 1; k < numX.size(); ++k) startX[k] = startX[k - 1] + numX[k - 1];
    }
    bool active(const Eng& e) const { return e.numTS >= start_ts; }
    void pos3(unsigned* p, unsigned l, unsigned i, unsigned j) const
    {
        p[ny] = l;
        p[nyP] = i;
        p[nyPP] = j;
    }
    void preV(Eng& e, unsigned tid) override
    {
        if (!active(e) || tid >= numX.size()) return;
        unsigned p[3], s[3];
        for (unsigned i = startX[tid]; i < startX[tid] + numX[tid]; ++i)
            for (unsigned j = 0; j < nl[1]; ++j) {
                pos3(p, line, i, j);
                pos3(s, line_shift, i, j);
                size_t k = (size_t)i * nl[1] + j;
                mP[k] = (T)e.GetVolt(nyP, s[0], s[1], s[2]) - cP[k] * (T)e.GetVolt(nyP, p[0], p[1], p[2]);
                mPP[k] = (T)e.GetVolt(nyPP, s[0], s[1], s[2]) - cPP[k] * (T)e.GetVolt(nyPP, p[0], p[1], p[2]);
            }
    }
    void postV(Eng& e, unsigned tid) override
    {
        if (!active(e) || tid >= numX.size()) return;
        unsigned s[3];
        for (unsigned i = startX[tid]; i < startX[tid] + numX[tid]; ++i)
            for (unsigned j = 0; j < nl[1]; ++j) {
                pos3(s, line_shift, i, j);
                size_t k = (size_t)i * nl[1] + j;
                mP[k] += cP[k] * (T)e.GetVolt(nyP, s[0], s[1], s[2]);
                mPP[k] += cPP[k] * (T)e.GetVolt(nyPP, s[0], s[1], s[2]);
            }
    }
    void applyV(Eng& e, unsigned tid) override
    {
        if (!active(e) || tid >= numX.size()) return;
        unsigned p[3];
        for (unsigned i = startX[tid]; i < startX[tid] + numX[tid]; ++i)
            for (unsigned j = 0; j < nl[1]; ++j) {
                pos3(p, line, i, j);
                size_t k = (size_t)i * nl[1] + j;
                e.SetVolt(nyP, p[0], p[1], p[2], mP[k]);
                e.SetVolt(nyPP, p[0], p[1], p[2], mPP[k]);
            }
    }
};

// --- lumped series / parallel RLC (ADE) ------------------------------------------
// Engine_Ext_LumpedRLC::DoPreVoltageUpdates/Apply2Voltages
// (FDTD/extensions/engine_ext_lumpedRLC.cpp:83-147); single-threaded hooks.
template <typename T, typename Eng>
struct RlcExt : ExtT<T, Eng> {
    unsigned count = 0;
    std::vector<unsigned> pos[3];
    std::vector<int> dir;
    std::vector<T> i2v, ilv, vvd, vv2, vj1, vj2, ib0, b1, b2;
    std::vector<T> hist[6];   // Vd[0..2], J[0..2] storage; rotated through the pointers below
    int vd[3] = {0, 1, 2}, jn[3] = {3, 4, 5};
    std::vector<T> Il;
    RlcExt() { this->prio = 0; }
    void preV(Eng&, unsigned tid) override
    {
        if (tid) return;
        int t = vd[2];
        vd[2] = vd[1];
        vd[1] = vd[0];
        vd[0] = t;
        for (unsigned k = 0; k < count; ++k) Il[k] += (i2v[k] * ilv[k]) * hist[vd[1]][k];
    }
    void applyV(Eng& e, unsigned tid) override
    {
        if (tid) return;
        int t = jn[2];
        jn[2] = jn[1];
        jn[1] = jn[0];
        jn[0] = t;
        std::vector<T>& V0 = hist[vd[0]];
        std::vector<T>& V2 = hist[vd[2]];
        std::vector<T>& J0 = hist[jn[0]];
        std::vector<T>& J1 = hist[jn[1]];
        std::vector<T>& J2 = hist[jn[2]];
        for (unsigned k = 0; k < count; ++k) V0[k] = (T)e.GetVolt(dir[k], pos[0][k], pos[1][k], pos[2][k]);
        for (unsigned k = 0; k < count; ++k) {
            V0[k] = vvd[k] * (V0[k] - Il[k] + vv2[k] * V2[k] + vj1[k] * J1[k] + vj2[k] * J2[k]);
            J0[k] = ib0[k] * (V0[k] - V2[k]) - (b1[k] * ib0[k]) * J1[k] - (b2[k] * ib0[k]) * J2[k];
        }
        for (unsigned k = 0; k < count; ++k) e.SetVolt(dir[k], pos[0][k], pos[1][k], pos[2][k], V0[k]);
    }
};

// --- Lorentz / Drude ADE --------------------------------------------------------------
// Engine_Ext_LorentzMaterial / Engine_Ext_Dispersive
template <typename T, typename Eng>
struct LorentzExt : ExtT<T, Eng> {
    unsigned order = 0;
    std::vector<unsigned> count;                 // per order
    std::vector<std::vector<unsigned>> pos[3];   // [dim][order][k]
    std::vector<std::vector<T>> v_int[3], v_ext[3], i_int[3], i_ext[3];
    std::vector<std::vector<T>> v_Lor[3], i_Lor[3];   // (omega^2 term), empty for Drude
    std::vector<std::vector<T>> volt_ADE[3], curr_ADE[3], volt_Lor[3], curr_Lor[3];
    std::vector<int> has_lorentz, has_volt, has_curr;
    LorentzExt() { this->prio = 0; }
    template <bool CURR>
    void pre(Eng& e)
    {
        for (unsigned o = 0; o < order; ++o) {
            if (CURR ? !has_curr[o] : !has_volt[o]) continue;
            for (unsigned k = 0; k < count[o]; ++k) {
                unsigned x = pos[0][o][k], y = pos[1][o][k], z = pos[2][o][k];
                for (int n = 0; n < 3; ++n) {
                    auto& ade = CURR ? curr_ADE[n][o] : volt_ADE[n][o];
                    auto& lor = CURR ? curr_Lor[n][o] : volt_Lor[n][o];
                    const auto& cint = CURR ? i_int[n][o] : v_int[n][o];
                    const auto& cext = CURR ? i_ext[n][o] : v_ext[n][o];
                    T f = CURR ? (T)e.GetCurr(n, x, y, z) : (T)e.GetVolt(n, x, y, z);
                    if (has_lorentz[o] & (CURR ? 2 : 1)) {
                        // engine_ext_lorentzmaterial.cpp:89-97: Lor first, then ADE
                        const auto& clor = CURR ? i_Lor[n][o] : v_Lor[n][o];
                        lor[k] += clor[k] * ade[k];
                        ade[k] *= cint[k];
                        ade[k] += cext[k] * (f - lor[k]);
                    } else {
                        ade[k] *= cint[k];
                        ade[k] += cext[k] * f;
                    }
                }
            }
        }
    }
    template <bool CURR>
    void apply(Eng& e)
    {
        for (unsigned o = 0; o < order; ++o) {
            if (CURR ? !has_curr[o] : !has_volt[o]) continue;
            for (unsigned k = 0; k < count[o]; ++k) {
                unsigned x = pos[0][o][k], y = pos[1][o][k], z = pos[2][o][k];
                for (int n = 0; n < 3; ++n) {
                    auto& ade = CURR ? curr_ADE[n][o] : volt_ADE[n][o];
                    if (CURR)
                        e.SetCurr(n, x, y, z, (T)e.GetCurr(n, x, y, z) - ade[k]);
                    else
                        e.SetVolt(n, x, y, z, (T)e.GetVolt(n, x, y, z) - ade[k]);
                }
            }
        }
    }
    void preV(Eng& e, unsigned tid) override { if (!tid) pre<false>(e); }
    void applyV(Eng& e, unsigned tid) override { if (!tid) apply<false>(e); }
    void preI(Eng& e, unsigned tid) override { if (!tid) pre<true>(e); }
    void applyI(Eng& e, unsigned tid) override { if (!tid) apply<true>(e); }
};

// --- steady-state detection (Engine_Ext_SteadyState::Apply2Voltages,
// FDTD/extensions/engine_ext_steadystate.cpp:50-107; priority +2e6, so its
// Apply2Voltages runs before every other apply hook).  The energy is
// Engine_Interface_FDTD::CalcFastEnergy (engine_interface_fdtd.cpp:302-347);
// abs() of the relative energy change is taken as fabs.
double fast_energy(const EngineBase& e);

template <typename T, typename Eng>
struct SteadyExt : ExtT<T, Eng> {
    unsigned p = 0;                                // m_TS_period
    std::vector<int> dir;
    std::vector<unsigned> pos[3];
    std::vector<std::vector<double>> rec;          // m_E_records, [probe][2p]
    double last_max_diff = 1, last_total_energy = 0;
    std::vector<unsigned> hist_ts;                 // timestep of every evaluation
    std::vector<double> hist_diff;                 // m_last_max_diff after it
    void applyV(Eng& e, unsigned tid) override
    {
        if (tid) return;
        const unsigned TS = e.numTS;
        const unsigned rel_pos = TS % (2 * p);
        for (size_t n = 0; n < rec.size(); ++n) rec[n][rel_pos] = e.getV(dir[n], pos[0][n], pos[1][n], pos[2][n]);
        if ((TS % p == 0) && (TS >= 2 * p)) {
            bool no_valid = true;
            last_max_diff = 0;
            const double curr_total_energy = fast_energy(e);
            if (last_total_energy > 0) {
                last_max_diff = std::fabs(curr_total_energy - last_total_energy) / last_total_energy;
                no_valid = false;
            }
            last_total_energy = curr_total_energy;
            unsigned old_pos = 0, new_pos = p;
            if (rel_pos <= p) {
                new_pos = 0;
                old_pos = p;
            }
            std::vector<double> curr_pow(rec.size()), diff_pow(rec.size());
            double max_pow = 0;
            for (size_t n = 0; n < rec.size(); ++n) {
                const double* buf = rec[n].data();
                curr_pow[n] = 0;
                diff_pow[n] = 0;
                for (unsigned nt = 0; nt < p; ++nt) {
                    curr_pow[n] += buf[nt + new_pos] * buf[nt + new_pos];
                    diff_pow[n] += (buf[nt + old_pos] - buf[nt + new_pos]) * (buf[nt + old_pos] - buf[nt + new_pos]);
                }
                max_pow = std::max(max_pow, curr_pow[n]);
            }
            for (size_t n = 0; n < rec.size(); ++n)
                if (curr_pow[n] > max_pow * 1e-2) {
                    last_max_diff = std::max(last_max_diff, diff_pow[n] / curr_pow[n]);
                    no_valid = false;
                }
            if (no_valid || last_max_diff > 1) last_max_diff = 1;
            hist_ts.push_back(TS);
            hist_diff.push_back(last_max_diff);
        }
    }
};

// ---------------------------------------------------------------------------
// extension list ordering (Engine::SortExtensionByPriority, FDTD/engine.cpp:87-98)
// ---------------------------------------------------------------------------
template <typename T, typename Eng>
static void sort_exts(std::vector<ExtT<T, Eng>*>& exts)
{
    std::stable_sort(exts.begin(), exts.end(),
                     [](ExtT<T, Eng>* a, ExtT<T, Eng>* b) { return a->prio < b->prio; });
    std::reverse(exts.begin(), exts.end());
}

// ---------------------------------------------------------------------------
// basic engine (N-I-J-K, z fastest), FDTD/engine.cpp
// ---------------------------------------------------------------------------
template <typename T>
struct BasicEngine : EngineBase {
    std::vector<T> volt, curr, vv, vi, ii, iv;
    std::vector<ExtT<T, BasicEngine<T>>*> exts;
    inline size_t idx(int n, unsigned x, unsigned y, unsigned z) const
    {
        return (((size_t)n * N[0] + x) * N[1] + y) * N[2] + z;
    }
    T GetVolt(int n, unsigned x, unsigned y, unsigned z) const { return volt[idx(n, x, y, z)]; }
    T GetCurr(int n, unsigned x, unsigned y, unsigned z) const { return curr[idx(n, x, y, z)]; }
    void SetVolt(int n, unsigned x, unsigned y, unsigned z, T v) { volt[idx(n, x, y, z)] = v; }
    void SetCurr(int n, unsigned x, unsigned y, unsigned z, T v) { curr[idx(n, x, y, z)] = v; }
    double getV(int n, unsigned x, unsigned y, unsigned z) const override { return GetVolt(n, x, y, z); }
    double getI(int n, unsigned x, unsigned y, unsigned z) const override { return GetCurr(n, x, y, z); }

    void UpdateVoltages(unsigned startX, unsigned numX)
    {
        const unsigned Ny = N[1], Nz = N[2];
        for (unsigned x = startX; x < startX + numX; ++x) {
            unsigned sx = x ? 1 : 0;
            for (unsigned y = 0; y < Ny; ++y) {
                unsigned sy = y ? 1 : 0;
                for (unsigned z = 0; z < Nz; ++z) {
                    unsigned sz = z ? 1 : 0;
                    size_t i0 = idx(0, x, y, z), i1 = idx(1, x, y, z), i2 = idx(2, x, y, z);
                    volt[i0] *= vv[i0];
                    volt[i0] += vi[i0] * (curr[idx(2, x, y, z)] - curr[idx(2, x, y - sy, z)] - curr[idx(1, x, y, z)] + curr[idx(1, x, y, z - sz)]);
                    volt[i1] *= vv[i1];
                    volt[i1] += vi[i1] * (curr[idx(0, x, y, z)] - curr[idx(0, x, y, z - sz)] - curr[idx(2, x, y, z)] + curr[idx(2, x - sx, y, z)]);
                    volt[i2] *= vv[i2];
                    volt[i2] += vi[i2] * (curr[idx(1, x, y, z)] - curr[idx(1, x - sx, y, z)] - curr[idx(0, x, y, z)] + curr[idx(0, x, y - sy, z)]);
                }
            }
        }
    }
    void UpdateCurrents(unsigned startX, unsigned numX)
    {
        const unsigned Ny = N[1], Nz = N[2];
        for (unsigned x = startX; x < startX + numX; ++x)
            for (unsigned y = 0; y < Ny - 1; ++y)
                for (unsigned z = 0; z < Nz - 1; ++z) {
                    size_t i0 = idx(0, x, y, z), i1 = idx(1, x, y, z), i2 = idx(2, x, y, z);
                    curr[i0] *= ii[i0];
                    curr[i0] += iv[i0] * (volt[idx(2, x, y, z)] - volt[idx(2, x, y + 1, z)] - volt[idx(1, x, y, z)] + volt[idx(1, x, y, z + 1)]);
                    curr[i1] *= ii[i1];
                    curr[i1] += iv[i1] * (volt[idx(0, x, y, z)] - volt[idx(0, x, y, z + 1)] - volt[idx(2, x, y, z)] + volt[idx(2, x + 1, y, z)]);
                    curr[i2] *= ii[i2];
                    curr[i2] += iv[i2] * (volt[idx(1, x, y, z)] - volt[idx(1, x + 1, y, z)] - volt[idx(0, x, y, z)] + volt[idx(0, x, y + 1, z)]);
                }
    }
    void iterate(unsigned steps) override
    {
        for (unsigned it = 0; it < steps; ++it) {
            for (int n = (int)exts.size() - 1; n >= 0; --n) exts[n]->preV(*this, 0);
            UpdateVoltages(0, N[0]);
            for (auto* e : exts) e->postV(*this, 0);
            for (auto* e : exts) e->applyV(*this, 0);
            for (int n = (int)exts.size() - 1; n >= 0; --n) exts[n]->preI(*this, 0);
            UpdateCurrents(0, N[0] - 1);
            for (auto* e : exts) e->postI(*this, 0);
            for (auto* e : exts) e->applyI(*this, 0);
            ++numTS;
        }
    }
};

// ---------------------------------------------------------------------------
// SSE engines: f4 lanes along z, lane l of vector v holds z = l*nV + v,
// I-J-K-N storage (component fastest)
// ---------------------------------------------------------------------------
struct SSEEngine : EngineBase {
    // flush-to-zero / denormals-are-zero as the reference's SSE engines run
    // (FDTD/engine_sse.cpp:43); off, the engine keeps denormals like the basic one
    bool ftz = true;
    unsigned nV;
    std::vector<f4vector> volt, curr;
    // plain coefficients (sse engine) or compressed table + index (sse-compressed)
    std::vector<f4vector> vv, vi, ii, iv;
    bool compressed = false;
    std::vector<f4vector> cvv[3], cvi[3], civ[3], cii[3];
    std::vector<uint32_t> opidx;  // [Nx][Ny][Nz] (only z < nV used), first-seen order
    std::vector<ExtT<float, SSEEngine>*> exts;
    // multithreading
    unsigned nthreads = 1;
    std::vector<std::thread> workers;
    Barrier* iterBarrier = nullptr;
    Barrier* startBarrier = nullptr;
    Barrier* stopBarrier = nullptr;
    std::atomic<bool> stop{false};
    unsigned iterTS = 0;
    std::vector<unsigned> tstart, tstop;

    inline size_t fidx(int n, unsigned x, unsigned y, unsigned v) const
    {
        return (size_t)n + 3 * ((size_t)v + nV * ((size_t)y + (size_t)N[1] * x));
    }
    float GetVolt(int n, unsigned x, unsigned y, unsigned z) const { return volt[fidx(n, x, y, z % nV)].f[z / nV]; }
    float GetCurr(int n, unsigned x, unsigned y, unsigned z) const { return curr[fidx(n, x, y, z % nV)].f[z / nV]; }
    void SetVolt(int n, unsigned x, unsigned y, unsigned z, float v) { volt[fidx(n, x, y, z % nV)].f[z / nV] = v; }
    void SetCurr(int n, unsigned x, unsigned y, unsigned z, float v) { curr[fidx(n, x, y, z % nV)].f[z / nV] = v; }
    double getV(int n, unsigned x, unsigned y, unsigned z) const override { return GetVolt(n, x, y, z); }
    double getI(int n, unsigned x, unsigned y, unsigned z) const override { return GetCurr(n, x, y, z); }

    void load_coeffs(const float* v_v, const float* v_i, const float* i_i, const float* i_v)
    {
        // Operator_sse::InitOperator + copy of the N-I-J-K operator into lanes
        size_t tot = (size_t)3 * N[0] * N[1] * nV;
        f4vector zero;
        zero.v = (v4sf){0, 0, 0, 0};
        vv.assign(tot, zero);
        vi.assign(tot, zero);
        ii.assign(tot, zero);
        iv.assign(tot, zero);
        for (int n = 0; n < 3; ++n)
            for (unsigned x = 0; x < N[0]; ++x)
                for (unsigned y = 0; y < N[1]; ++y)
                    for (unsigned z = 0; z < N[2]; ++z) {
                        size_t s = (((size_t)n * N[0] + x) * N[1] + y) * N[2] + z;
                        size_t d = fidx(n, x, y, z % nV);
                        vv[d].f[z / nV] = v_v[s];
                        vi[d].f[z / nV] = v_i[s];
                        ii[d].f[z / nV] = i_i[s];
                        iv[d].f[z / nV] = i_v[s];
                    }
    }

    void compress()
    {
        // Operator_SSE_Compressed::CompressOperator; std::map keyed by the
        // float-value ordering of SSE_coeff::operator< (lexicographic over
        // n, lane, vv/vi/iv/ii)
        struct Key {
            float c[48];
            bool operator<(const Key& o) const
            {
                for (int i = 0; i < 48; ++i) {
                    if (c[i] > o.c[i]) return false;
                    if (c[i] < o.c[i]) return true;
                }
                return false;
            }
        };
        std::map<Key, uint32_t> lut;
        opidx.assign((size_t)N[0] * N[1] * N[2], 0);
        for (int n = 0; n < 3; ++n) {
            cvv[n].clear(); cvi[n].clear(); civ[n].clear(); cii[n].clear();
        }
        for (unsigned x = 0; x < N[0]; ++x)
            for (unsigned y = 0; y < N[1]; ++y)
                for (unsigned v = 0; v < nV; ++v) {
                    Key k;
                    int q = 0;
                    for (int n = 0; n < 3; ++n)
                        for (int c = 0; c < 4; ++c) {
                            size_t d = fidx(n, x, y, v);
                            k.c[q++] = vv[d].f[c];
                            k.c[q++] = vi[d].f[c];
                            k.c[q++] = iv[d].f[c];
                            k.c[q++] = ii[d].f[c];
                        }
                    auto it = lut.find(k);
                    uint32_t index;
                    if (it == lut.end()) {
                        index = (uint32_t)cvv[0].size();
                        for (int n = 0; n < 3; ++n) {
                            size_t d = fidx(n, x, y, v);
                            cvv[n].push_back(vv[d]);
                            cvi[n].push_back(vi[d]);
                            civ[n].push_back(iv[d]);
                            cii[n].push_back(ii[d]);
                        }
                        lut[k] = index;
                    } else
                        index = it->second;
                    opidx[((size_t)x * N[1] + y) * N[2] + v] = index;
                }
        compressed = true;
        vv.clear(); vi.clear(); ii.clear(); iv.clear();
        vv.shrink_to_fit(); vi.shrink_to_fit(); ii.shrink_to_fit(); iv.shrink_to_fit();
    }

    inline const f4vector& CVV(int n, unsigned x, unsigned y, unsigned v) const
    {
        return compressed ? cvv[n][opidx[((size_t)x * N[1] + y) * N[2] + v]] : vv[fidx(n, x, y, v)];
    }
    inline const f4vector& CVI(int n, unsigned x, unsigned y, unsigned v) const
    {
        return compressed ? cvi[n][opidx[((size_t)x * N[1] + y) * N[2] + v]] : vi[fidx(n, x, y, v)];
    }
    inline const f4vector& CII(int n, unsigned x, unsigned y, unsigned v) const
    {
        return compressed ? cii[n][opidx[((size_t)x * N[1] + y) * N[2] + v]] : ii[fidx(n, x, y, v)];
    }
    inline const f4vector& CIV(int n, unsigned x, unsigned y, unsigned v) const
    {
        return compressed ? civ[n][opidx[((size_t)x * N[1] + y) * N[2] + v]] : iv[fidx(n, x, y, v)];
    }

    // Engine_sse::UpdateVoltages / Engine_SSE_Compressed::UpdateVoltages
    void UpdateVoltages(unsigned startX, unsigned numX)
    {
        f4vector temp;
        for (unsigned x = startX; x < startX + numX; ++x) {
            unsigned sx = x ? 1 : 0;
            for (unsigned y = 0; y < N[1]; ++y) {
                unsigned sy = y ? 1 : 0;
                for (unsigned v = 1; v < nV; ++v) {
                    f4vector& V0 = volt[fidx(0, x, y, v)];
                    V0.v *= CVV(0, x, y, v).v;
                    V0.v += CVI(0, x, y, v).v * (curr[fidx(2, x, y, v)].v - curr[fidx(2, x, y - sy, v)].v - curr[fidx(1, x, y, v)].v + curr[fidx(1, x, y, v - 1)].v);
                    f4vector& V1 = volt[fidx(1, x, y, v)];
                    V1.v *= CVV(1, x, y, v).v;
                    V1.v += CVI(1, x, y, v).v * (curr[fidx(0, x, y, v)].v - curr[fidx(0, x, y, v - 1)].v - curr[fidx(2, x, y, v)].v + curr[fidx(2, x - sx, y, v)].v);
                    f4vector& V2 = volt[fidx(2, x, y, v)];
                    V2.v *= CVV(2, x, y, v).v;
                    V2.v += CVI(2, x, y, v).v * (curr[fidx(1, x, y, v)].v - curr[fidx(1, x - sx, y, v)].v - curr[fidx(0, x, y, v)].v + curr[fidx(0, x, y - sy, v)].v);
                }
                // v = 0: the z-1 neighbour comes from the previous lane of the last vector, lane 0 gets 0
                const f4vector& c1e = curr[fidx(1, x, y, nV - 1)];
                temp.f[0] = 0; temp.f[1] = c1e.f[0]; temp.f[2] = c1e.f[1]; temp.f[3] = c1e.f[2];
                f4vector& V0 = volt[fidx(0, x, y, 0)];
                V0.v *= CVV(0, x, y, 0).v;
                V0.v += CVI(0, x, y, 0).v * (curr[fidx(2, x, y, 0)].v - curr[fidx(2, x, y - sy, 0)].v - curr[fidx(1, x, y, 0)].v + temp.v);
                const f4vector& c0e = curr[fidx(0, x, y, nV - 1)];
                temp.f[0] = 0; temp.f[1] = c0e.f[0]; temp.f[2] = c0e.f[1]; temp.f[3] = c0e.f[2];
                f4vector& V1 = volt[fidx(1, x, y, 0)];
                V1.v *= CVV(1, x, y, 0).v;
                V1.v += CVI(1, x, y, 0).v * (curr[fidx(0, x, y, 0)].v - temp.v - curr[fidx(2, x, y, 0)].v + curr[fidx(2, x - sx, y, 0)].v);
                f4vector& V2 = volt[fidx(2, x, y, 0)];
                V2.v *= CVV(2, x, y, 0).v;
                V2.v += CVI(2, x, y, 0).v * (curr[fidx(1, x, y, 0)].v - curr[fidx(1, x - sx, y, 0)].v - curr[fidx(0, x, y, 0)].v + curr[fidx(0, x, y - sy, 0)].v);
            }
        }
    }

    // Engine_sse::UpdateCurrents / Engine_SSE_Compressed::UpdateCurrents
    void UpdateCurrents(unsigned startX, unsigned numX)
    {
        f4vector temp;
        for (unsigned x = startX; x < startX + numX; ++x)
            for (unsigned y = 0; y < N[1] - 1; ++y) {
                for (unsigned v = 0; v + 1 < nV; ++v) {
                    f4vector& I0 = curr[fidx(0, x, y, v)];
                    I0.v *= CII(0, x, y, v).v;
                    I0.v += CIV(0, x, y, v).v * (volt[fidx(2, x, y, v)].v - volt[fidx(2, x, y + 1, v)].v - volt[fidx(1, x, y, v)].v + volt[fidx(1, x, y, v + 1)].v);
                    f4vector& I1 = curr[fidx(1, x, y, v)];
                    I1.v *= CII(1, x, y, v).v;
                    I1.v += CIV(1, x, y, v).v * (volt[fidx(0, x, y, v)].v - volt[fidx(0, x, y, v + 1)].v - volt[fidx(2, x, y, v)].v + volt[fidx(2, x + 1, y, v)].v);
                    f4vector& I2 = curr[fidx(2, x, y, v)];
                    I2.v *= CII(2, x, y, v).v;
                    I2.v += CIV(2, x, y, v).v * (volt[fidx(1, x, y, v)].v - volt[fidx(1, x + 1, y, v)].v - volt[fidx(0, x, y, v)].v + volt[fidx(0, x, y + 1, v)].v);
                }
                unsigned v = nV - 1;
                const f4vector& v10 = volt[fidx(1, x, y, 0)];
                temp.f[0] = v10.f[1]; temp.f[1] = v10.f[2]; temp.f[2] = v10.f[3]; temp.f[3] = 0;
                f4vector& I0 = curr[fidx(0, x, y, v)];
                I0.v *= CII(0, x, y, v).v;
                I0.v += CIV(0, x, y, v).v * (volt[fidx(2, x, y, v)].v - volt[fidx(2, x, y + 1, v)].v - volt[fidx(1, x, y, v)].v + temp.v);
                const f4vector& v00 = volt[fidx(0, x, y, 0)];
                temp.f[0] = v00.f[1]; temp.f[1] = v00.f[2]; temp.f[2] = v00.f[3]; temp.f[3] = 0;
                f4vector& I1 = curr[fidx(1, x, y, v)];
                I1.v *= CII(1, x, y, v).v;
                I1.v += CIV(1, x, y, v).v * (volt[fidx(0, x, y, v)].v - temp.v - volt[fidx(2, x, y, v)].v + volt[fidx(2, x + 1, y, v)].v);
                f4vector& I2 = curr[fidx(2, x, y, v)];
                I2.v *= CII(2, x, y, v).v;
                I2.v += CIV(2, x, y, v).v * (volt[fidx(1, x, y, v)].v - volt[fidx(1, x + 1, y, v)].v - volt[fidx(0, x, y, v)].v + volt[fidx(0, x, y + 1, v)].v);
            }
    }

    // --- hooks with a barrier after every extension (engine_multithread.cpp:234-263)
    void hooks(int phase, unsigned tid)
    {
        const int m = (int)exts.size();
        if (phase == 0 || phase == 3) {
            for (int n = m - 1; n >= 0; --n) {
                if (phase == 0) exts[n]->preV(*this, tid); else exts[n]->preI(*this, tid);
                if (nthreads > 1) iterBarrier->wait();
            }
        } else {
            for (int n = 0; n < m; ++n) {
                switch (phase) {
                case 1: exts[n]->postV(*this, tid); break;
                case 2: exts[n]->applyV(*this, tid); break;
                case 4: exts[n]->postI(*this, tid); break;
                case 5: exts[n]->applyI(*this, tid); break;
                }
                if (nthreads > 1) iterBarrier->wait();
            }
        }
    }

    void worker(unsigned tid)
    {
        if (ftz) {
            _MM_SET_FLUSH_ZERO_MODE(_MM_FLUSH_ZERO_ON);      // Denormal::Disable (tools/denormal.h)
            _MM_SET_DENORMALS_ZERO_MODE(_MM_DENORMALS_ZERO_ON);
        }
        unsigned start = tstart[tid], stopx = tstop[tid];
        unsigned stop_h = (tid == nthreads - 1) ? stopx - 1 : stopx;
        while (true) {
            startBarrier->wait();
            if (stop.load()) return;
            for (unsigned it = 0; it < iterTS; ++it) {
                hooks(0, tid);
                UpdateVoltages(start, stopx - start + 1);
                iterBarrier->wait();
                hooks(1, tid);
                hooks(2, tid);
                hooks(3, tid);
                UpdateCurrents(start, stop_h - start + 1);
                iterBarrier->wait();
                hooks(4, tid);
                hooks(5, tid);
                if (tid == 0) ++numTS;
            }
            stopBarrier->wait();
        }
    }

    void start_threads(unsigned n)
    {
        nthreads = n;
        // Operator_Multithread::CalcStartStopLines (operator_multithread.cpp:93-114)
        std::vector<unsigned> jpt = assign_jobs(N[0], n);
        unsigned line = 0;
        for (unsigned k = 0; k < n; ++k) {
            tstart.push_back(line);
            tstop.push_back(line + jpt[k] - 1);
            line += jpt[k];
        }
        iterBarrier = new Barrier(n);
        startBarrier = new Barrier(n + 1);
        stopBarrier = new Barrier(n + 1);
        for (auto* e : exts) e->setThreads(n);
        for (unsigned k = 0; k < n; ++k) workers.emplace_back(&SSEEngine::worker, this, k);
    }

    void iterate(unsigned steps) override
    {
        if (nthreads > 1 || !workers.empty()) {
            iterTS = steps;
            startBarrier->wait();
            stopBarrier->wait();
            return;
        }
        unsigned csr = _mm_getcsr();
        if (ftz) {
            _MM_SET_FLUSH_ZERO_MODE(_MM_FLUSH_ZERO_ON);
            _MM_SET_DENORMALS_ZERO_MODE(_MM_DENORMALS_ZERO_ON);
        }
        for (unsigned it = 0; it < steps; ++it) {
            hooks(0, 0);
            UpdateVoltages(0, N[0]);
            hooks(1, 0);
            hooks(2, 0);
            hooks(3, 0);
            UpdateCurrents(0, N[0] - 1);
            hooks(4, 0);
            hooks(5, 0);
            ++numTS;
        }
        _mm_setcsr(csr);
    }

    ~SSEEngine()
    {
        if (!workers.empty()) {
            stop.store(true);
            startBarrier->wait();
            for (auto& t : workers) t.join();
        }
        delete iterBarrier;
        delete startBarrier;
        delete stopBarrier;
        for (auto* e : exts) delete e;
    }
};

// ---------------------------------------------------------------------------
// handle + C ABI for ctypes
// ---------------------------------------------------------------------------
enum Kind { K_BASIC_F32 = 0, K_BASIC_F64 = 1, K_SSE = 2, K_SSE_COMPRESSED = 3, K_MT = 4 };

struct Handle {
    int kind;
    EngineBase* eng = nullptr;
    BasicEngine<float>* bf = nullptr;
    BasicEngine<double>* bd = nullptr;
    SSEEngine* se = nullptr;
    unsigned nthreads = 1;
    bool inited = false;
    void* ss = nullptr;   // steady-state extension (owned by the engine's extension list)
};

template <typename Eng, typename T>
static void add_common(std::vector<ExtT<T, Eng>*>& exts, ExtT<T, Eng>* e) { exts.push_back(e); }

template <typename T, typename Eng>
static ExcExt<T, Eng>* make_exc(const float* sigV, const float* sigC, unsigned length, int period_steps,
                                unsigned nv, const unsigned* vidx, const unsigned short* vdir, const float* vamp, const unsigned* vdelay,
                                unsigned nc, const unsigned* cidx, const unsigned short* cdir, const float* camp, const unsigned* cdelay)
{
    auto* e = new ExcExt<T, Eng>();
    e->sigV.assign(sigV, sigV + length);
    e->sigC.assign(sigC, sigC + length);
    e->length = length;
    e->period_steps = period_steps;
    for (int d = 0; d < 3; ++d) {
        e->vidx[d].assign(vidx + (size_t)d * nv, vidx + (size_t)(d + 1) * nv);
        e->cidx[d].assign(cidx + (size_t)d * nc, cidx + (size_t)(d + 1) * nc);
    }
    e->vdir.assign(vdir, vdir + nv);
    e->vamp.assign(vamp, vamp + nv);
    e->vdelay.assign(vdelay, vdelay + nv);
    e->cdir.assign(cdir, cdir + nc);
    e->camp.assign(camp, camp + nc);
    e->cdelay.assign(cdelay, cdelay + nc);
    return e;
}

template <typename T, typename Eng>
static UpmlExt<T, Eng>* make_upml(const unsigned* start, const unsigned* num, const float* const* arr)
{
    auto* u = new UpmlExt<T, Eng>();
    size_t tot = (size_t)3 * num[0] * num[1] * num[2];
    for (int k = 0; k < 3; ++k) {
        u->start[k] = start[k];
        u->num[k] = num[k];
    }
    std::vector<T>* dst[6] = {&u->vv, &u->vvfo, &u->vvfn, &u->ii, &u->iifo, &u->iifn};
    for (int k = 0; k < 6; ++k) dst[k]->assign(arr[k], arr[k] + tot);
    u->vflux.assign(tot, 0);
    u->cflux.assign(tot, 0);
    u->setThreads(1);
    return u;
}

template <typename T, typename Eng>
static MurExt<T, Eng>* make_mur(const unsigned* N, int ny, unsigned line, unsigned shift, const float* cP, const float* cPP, unsigned start_ts)
{
    auto* m = new MurExt<T, Eng>();
    m->ny = ny;
    m->nyP = (ny + 1) % 3;
    m->nyPP = (ny + 2) % 3;
    m->line = line;
    m->line_shift = shift;
    m->nl[0] = N[m->nyP];
    m->nl[1] = N[m->nyPP];
    m->start_ts = start_ts;
    size_t tot = (size_t)m->nl[0] * m->nl[1];
    m->cP.assign(cP, cP + tot);
    m->cPP.assign(cPP, cPP + tot);
    m->mP.assign(tot, 0);
    m->mPP.assign(tot, 0);
    m->setThreads(1);
    return m;
}

extern "C" {

void* ref_create(int kind, const unsigned* N, const float* vv, const float* vi, const float* ii, const float* iv,
                 unsigned nthreads)
{
    Handle* h = new Handle();
    h->kind = kind;
    size_t tot = (size_t)3 * N[0] * N[1] * N[2];
    if (kind == K_BASIC_F32 || kind == K_BASIC_F64) {
        if (kind == K_BASIC_F32) {
            auto* e = new BasicEngine<float>();
            for (int k = 0; k < 3; ++k) e->N[k] = N[k];
            e->volt.assign(tot, 0); e->curr.assign(tot, 0);
            e->vv.assign(vv, vv + tot); e->vi.assign(vi, vi + tot);
            e->ii.assign(ii, ii + tot); e->iv.assign(iv, iv + tot);
            h->bf = e; h->eng = e;
        } else {
            auto* e = new BasicEngine<double>();
            for (int k = 0; k < 3; ++k) e->N[k] = N[k];
            e->volt.assign(tot, 0); e->curr.assign(tot, 0);
            e->vv.assign(vv, vv + tot); e->vi.assign(vi, vi + tot);
            e->ii.assign(ii, ii + tot); e->iv.assign(iv, iv + tot);
            h->bd = e; h->eng = e;
        }
    } else {
        auto* e = new SSEEngine();
        for (int k = 0; k < 3; ++k) e->N[k] = N[k];
        e->nV = (unsigned)std::ceil((double)N[2] / 4.0);
        f4vector zero;
        zero.v = (v4sf){0, 0, 0, 0};
        size_t ftot = (size_t)3 * N[0] * N[1] * e->nV;
        e->volt.assign(ftot, zero);
        e->curr.assign(ftot, zero);
        e->load_coeffs(vv, vi, ii, iv);
        if (kind == K_SSE_COMPRESSED || kind == K_MT) e->compress();
        h->se = e; h->eng = e;
        h->nthreads = (kind == K_MT) ? std::max(1u, nthreads) : 1;
    }
    return h;
}

// the SSE engines' flush-to-zero mode (set before the first iterate)
void ref_set_ftz(void* hh, int on)
{
    Handle* h = (Handle*)hh;
    if (h->se) h->se->ftz = on != 0;
}

// double-precision coefficients for the fp64 basic engine
void ref_set_coeffs_f64(void* hh, const double* vv, const double* vi, const double* ii, const double* iv)
{
    Handle* h = (Handle*)hh;
    if (!h->bd) return;
    size_t tot = h->bd->vv.size();
    h->bd->vv.assign(vv, vv + tot); h->bd->vi.assign(vi, vi + tot);
    h->bd->ii.assign(ii, ii + tot); h->bd->iv.assign(iv, iv + tot);
}

void ref_add_excitation(void* hh, const float* sigV, const float* sigC, unsigned length, int period_steps,
                        unsigned nv, const unsigned* vidx, const unsigned short* vdir, const float* vamp, const unsigned* vdelay,
                        unsigned nc, const unsigned* cidx, const unsigned short* cdir, const float* camp, const unsigned* cdelay)
{
    Handle* h = (Handle*)hh;
    if (h->bf) h->bf->exts.push_back(make_exc<float, BasicEngine<float>>(sigV, sigC, length, period_steps, nv, vidx, vdir, vamp, vdelay, nc, cidx, cdir, camp, cdelay));
    if (h->bd) h->bd->exts.push_back(make_exc<double, BasicEngine<double>>(sigV, sigC, length, period_steps, nv, vidx, vdir, vamp, vdelay, nc, cidx, cdir, camp, cdelay));
    if (h->se) h->se->exts.push_back(make_exc<float, SSEEngine>(sigV, sigC, length, period_steps, nv, vidx, vdir, vamp, vdelay, nc, cidx, cdir, camp, cdelay));
}

// arr: vv, vvfo, vvfn, ii, iifo, iifn, each [3][num0][num1][num2]
void ref_add_upml(void* hh, const unsigned* start, const unsigned* num, const float* vv, const float* vvfo,
                  const float* vvfn, const float* ii, const float* iifo, const float* iifn)
{
    Handle* h = (Handle*)hh;
    const float* arr[6] = {vv, vvfo, vvfn, ii, iifo, iifn};
    if (h->bf) h->bf->exts.push_back(make_upml<float, BasicEngine<float>>(start, num, arr));
    if (h->bd) h->bd->exts.push_back(make_upml<double, BasicEngine<double>>(start, num, arr));
    if (h->se) h->se->exts.push_back(make_upml<float, SSEEngine>(start, num, arr));
}

void ref_add_mur(void* hh, int ny, unsigned line, unsigned shift, const float* cP, const float* cPP, unsigned start_ts)
{
    Handle* h = (Handle*)hh;
    if (h->bf) h->bf->exts.push_back(make_mur<float, BasicEngine<float>>(h->bf->N, ny, line, shift, cP, cPP, start_ts));
    if (h->bd) h->bd->exts.push_back(make_mur<double, BasicEngine<double>>(h->bd->N, ny, line, shift, cP, cPP, start_ts));
    if (h->se) h->se->exts.push_back(make_mur<float, SSEEngine>(h->se->N, ny, line, shift, cP, cPP, start_ts));
}

// Lorentz/Drude ADE lists.  Per order o: count[o] cells, positions pos[3][count],
// coefficient arrays [3][count] (v_int, v_ext, v_lor, i_int, i_ext, i_lor);
// flags: has_lorentz (bit 0: voltage Lorentz pole, bit 1: current Lorentz pole,
// m_volt_Lor_ADE_On / m_curr_Lor_ADE_On), has_volt, has_curr.
void ref_add_lorentz(void* hh, unsigned order, const unsigned* count, const unsigned* const* pos,
                     const float* const* v_int, const float* const* v_ext, const float* const* v_lor,
                     const float* const* i_int, const float* const* i_ext, const float* const* i_lor,
                     const int* has_lorentz, const int* has_volt, const int* has_curr)
{
    Handle* h = (Handle*)hh;
    auto build = [&](auto* L) {
        L->order = order;
        L->count.assign(count, count + order);
        L->has_lorentz.assign(has_lorentz, has_lorentz + order);
        L->has_volt.assign(has_volt, has_volt + order);
        L->has_curr.assign(has_curr, has_curr + order);
        for (int d = 0; d < 3; ++d) {
            L->pos[d].resize(order);
            L->v_int[d].resize(order); L->v_ext[d].resize(order); L->v_Lor[d].resize(order);
            L->i_int[d].resize(order); L->i_ext[d].resize(order); L->i_Lor[d].resize(order);
            L->volt_ADE[d].resize(order); L->curr_ADE[d].resize(order);
            L->volt_Lor[d].resize(order); L->curr_Lor[d].resize(order);
        }
        for (unsigned o = 0; o < order; ++o) {
            unsigned c = count[o];
            for (int d = 0; d < 3; ++d) {
                L->pos[d][o].assign(pos[o] + (size_t)d * c, pos[o] + (size_t)(d + 1) * c);
                L->v_int[d][o].assign(v_int[o] + (size_t)d * c, v_int[o] + (size_t)(d + 1) * c);
                L->v_ext[d][o].assign(v_ext[o] + (size_t)d * c, v_ext[o] + (size_t)(d + 1) * c);
                L->i_int[d][o].assign(i_int[o] + (size_t)d * c, i_int[o] + (size_t)(d + 1) * c);
                L->i_ext[d][o].assign(i_ext[o] + (size_t)d * c, i_ext[o] + (size_t)(d + 1) * c);
                if (has_lorentz[o] & 1) L->v_Lor[d][o].assign(v_lor[o] + (size_t)d * c, v_lor[o] + (size_t)(d + 1) * c);
                if (has_lorentz[o] & 2) L->i_Lor[d][o].assign(i_lor[o] + (size_t)d * c, i_lor[o] + (size_t)(d + 1) * c);
                L->volt_ADE[d][o].assign(c, 0); L->curr_ADE[d][o].assign(c, 0);
                L->volt_Lor[d][o].assign(c, 0); L->curr_Lor[d][o].assign(c, 0);
            }
        }
    };
    if (h->bf) { auto* L = new LorentzExt<float, BasicEngine<float>>(); build(L); h->bf->exts.push_back(L); }
    if (h->bd) { auto* L = new LorentzExt<double, BasicEngine<double>>(); build(L); h->bd->exts.push_back(L); }
    if (h->se) { auto* L = new LorentzExt<float, SSEEngine>(); build(L); h->se->exts.push_back(L); }
}

// lumped RLC ADE list (Operator_Ext_LumpedRLC tables): pos[3][count], dir[count],
// coefficient arrays [count]: i2v, ilv, vvd, vv2, vj1, vj2, ib0, b1, b2
void ref_add_rlc(void* hh, unsigned count, const unsigned* pos, const int* dir, const float* const* c)
{
    Handle* h = (Handle*)hh;
    auto build = [&](auto* R) {
        R->count = count;
        for (int d = 0; d < 3; ++d) R->pos[d].assign(pos + (size_t)d * count, pos + (size_t)(d + 1) * count);
        R->dir.assign(dir, dir + count);
        decltype(&R->i2v) dst[9] = {&R->i2v, &R->ilv, &R->vvd, &R->vv2, &R->vj1, &R->vj2, &R->ib0, &R->b1, &R->b2};
        for (int q = 0; q < 9; ++q) dst[q]->assign(c[q], c[q] + count);
        for (int k = 0; k < 6; ++k) R->hist[k].assign(count, 0);
        R->Il.assign(count, 0);
    };
    if (h->bf) { auto* R = new RlcExt<float, BasicEngine<float>>(); build(R); h->bf->exts.push_back(R); }
    if (h->bd) { auto* R = new RlcExt<double, BasicEngine<double>>(); build(R); h->bd->exts.push_back(R); }
    if (h->se) { auto* R = new RlcExt<float, SSEEngine>(); build(R); h->se->exts.push_back(R); }
}

void ref_init(void* hh)
{
    Handle* h = (Handle*)hh;
    if (h->inited) return;
    if (h->bf) sort_exts(h->bf->exts);
    if (h->bd) sort_exts(h->bd->exts);
    if (h->se) {
        sort_exts(h->se->exts);
        if (h->kind == K_MT) h->se->start_threads(h->nthreads);
    }
    h->inited = true;
}

void ref_iterate(void* hh, unsigned steps)
{
    Handle* h = (Handle*)hh;
    ref_init(hh);
    h->eng->iterate(steps);
}

unsigned ref_numts(void* hh) { return ((Handle*)hh)->eng->numTS; }

// copy fields into N-I-J-K float64 buffers (for probes and comparisons)
void ref_get_fields(void* hh, double* volt, double* curr)
{
    Handle* h = (Handle*)hh;
    EngineBase* e = h->eng;
    size_t q = 0;
    for (int n = 0; n < 3; ++n)
        for (unsigned x = 0; x < e->N[0]; ++x)
            for (unsigned y = 0; y < e->N[1]; ++y)
                for (unsigned z = 0; z < e->N[2]; ++z, ++q) {
                    if (volt) volt[q] = e->getV(n, x, y, z);
                    if (curr) curr[q] = e->getI(n, x, y, z);
                }
}

// direct pointers to the basic engines' N-I-J-K storage (float or double)
void* ref_volt_ptr(void* hh)
{
    Handle* h = (Handle*)hh;
    if (h->bf) return h->bf->volt.data();
    if (h->bd) return h->bd->volt.data();
    return nullptr;
}
void* ref_curr_ptr(void* hh)
{
    Handle* h = (Handle*)hh;
    if (h->bf) return h->bf->curr.data();
    if (h->bd) return h->bd->curr.data();
    return nullptr;
}

unsigned ref_num_unique(void* hh)
{
    Handle* h = (Handle*)hh;
    return (h->se && h->se->compressed) ? (unsigned)h->se->cvv[0].size() : 0;
}

// compressed index in the reference's m_Op_index layout [Nx][Ny][nV]
void ref_get_opindex(void* hh, unsigned* out)
{
    Handle* h = (Handle*)hh;
    SSEEngine* e = h->se;
    if (!e || !e->compressed) return;
    size_t q = 0;
    for (unsigned x = 0; x < e->N[0]; ++x)
        for (unsigned y = 0; y < e->N[1]; ++y)
            for (unsigned v = 0; v < e->nV; ++v) out[q++] = e->opidx[((size_t)x * e->N[1] + y) * e->N[2] + v];
}

// compressed tables f4_{vv,vi,ii,iv}_Compressed[n] (which = 0..3), U x 4 floats
void ref_get_ctab(void* hh, int which, int n, float* out)
{
    Handle* h = (Handle*)hh;
    SSEEngine* e = h->se;
    if (!e || !e->compressed || which < 0 || which > 3 || n < 0 || n > 2) return;
    const std::vector<f4vector>* t[4] = {e->cvv, e->cvi, e->cii, e->civ};
    const auto& v = t[which][n];
    for (size_t u = 0; u < v.size(); ++u)
        for (int l = 0; l < 4; ++l) out[4 * u + l] = v[u].f[l];
}

// steady-state detection with period p (timesteps) and probes dir[n], pos[3][n]
void ref_add_steady_state(void* hh, unsigned p, unsigned n, const int* dir, const unsigned* pos)
{
    Handle* h = (Handle*)hh;
    auto build = [&](auto* S) {
        S->prio = 2e6;
        S->p = p;
        S->dir.assign(dir, dir + n);
        for (int d = 0; d < 3; ++d) S->pos[d].assign(pos + (size_t)d * n, pos + (size_t)(d + 1) * n);
        S->rec.assign(n, std::vector<double>(2 * (size_t)p, 0.0));
        h->ss = S;
    };
    if (h->bf) { auto* S = new SteadyExt<float, BasicEngine<float>>(); build(S); h->bf->exts.push_back(S); }
    else if (h->bd) { auto* S = new SteadyExt<double, BasicEngine<double>>(); build(S); h->bd->exts.push_back(S); }
    else if (h->se) { auto* S = new SteadyExt<float, SSEEngine>(); build(S); h->se->exts.push_back(S); }
}

// evaluations so far: timesteps and m_last_max_diff values (up to cap)
unsigned ref_steady_history(void* hh, unsigned* ts, double* diff, unsigned cap)
{
    Handle* h = (Handle*)hh;
    if (!h->ss) return 0;
    const std::vector<unsigned>* hts = nullptr;
    const std::vector<double>* hd = nullptr;
    if (h->bf) { auto* S = (SteadyExt<float, BasicEngine<float>>*)h->ss; hts = &S->hist_ts; hd = &S->hist_diff; }
    else if (h->bd) { auto* S = (SteadyExt<double, BasicEngine<double>>*)h->ss; hts = &S->hist_ts; hd = &S->hist_diff; }
    else { auto* S = (SteadyExt<float, SSEEngine>*)h->ss; hts = &S->hist_ts; hd = &S->hist_diff; }
    const unsigned n = (unsigned)std::min<size_t>(cap, hts->size());
    for (unsigned k = 0; k < n; ++k) {
        ts[k] = (*hts)[k];
        diff[k] = (*hd)[k];
    }
    return (unsigned)hts->size();
}

// CalcFastEnergy (engine_interface_fdtd.cpp:302-347): sum over [0,N-1)^3 in double
double ref_energy(void* hh) { return fast_energy(*((Handle*)hh)->eng); }

}  // extern "C"

double fast_energy(const EngineBase& eb)
{
    const EngineBase* e = &eb;
    double E = 0, H = 0;
    for (unsigned x = 0; x + 1 < e->N[0]; ++x)
        for (unsigned y = 0; y + 1 < e->N[1]; ++y)
            for (unsigned z = 0; z + 1 < e->N[2]; ++z)
                for (int n = 0; n < 3; ++n) {
                    double v = e->getV(n, x, y, z), c = e->getI(n, x, y, z);
                    E += v * v;
                    H += c * c;
                }
    return 8.85418781762e-12 * E + 1.256637062e-6 * H;
}

extern "C" {

void ref_destroy(void* hh)
{
    Handle* h = (Handle*)hh;
    if (h->bf) { for (auto* e : h->bf->exts) delete e; delete h->bf; }
    if (h->bd) { for (auto* e : h->bd->exts) delete e; delete h->bd; }
    delete h->se;
    delete h;
}

}  // extern "C"
