/*
 * ems_hip.h — MI355X (gfx950) EC-FDTD engine, C ABI.
 *
 * The drop-in boundary for openEMS's timestep loop.  An engine handle owns the
 * device copies of the operator, the fields and every engine extension, and
 * advances them with hand-written HIP kernels.  Host pointers passed in use
 * the reference's layouts (N-I-J-K float32 for fields and coefficients);
 * the device layout (structure of arrays, x fastest) stays internal.
 *
 * Reference interfaces replaced (paths relative to the openEMS tree):
 *   ems_create            Operator::CreateEngine / Engine::Init           FDTD/operator.cpp:68-72, FDTD/engine.cpp:51-59
 *   ems_create_layout     same, with the device coefficient layout chosen explicitly
 *   ems_create_sse_compressed
 *                         Operator_SSE_Compressed::CreateEngine on the compressed operator
 *                                                    FDTD/operator_sse_compressed.cpp:114-175, FDTD/operator_sse_compressed.h:49-77
 *   ems_get_layout        Operator_SSE_Compressed::ShowStat (compression state, unique count)
 *                                                    FDTD/operator_sse_compressed.cpp:105-112
 *   ems_compress_operator Operator_SSE_Compressed::CompressOperator, per cell (host only)
 *                                                    FDTD/operator_sse_compressed.cpp:114-175
 *   ems_expand_sse_compressed
 *                         Operator_SSE_Compressed::GetVV/GetVI/GetII/GetIV over the whole mesh (host only)
 *                                                    FDTD/operator_sse_compressed.h:49-52
 *   ems_set_excitation    Operator_Ext_Excitation -> Engine_Ext_Excitation FDTD/extensions/engine_ext_excitation.cpp:24-99
 *   ems_add_upml          Operator_Ext_UPML -> Engine_Ext_UPML            FDTD/extensions/engine_ext_upml.cpp:24-234
 *   ems_add_mur           Operator_Ext_Mur_ABC -> Engine_Ext_Mur_ABC      FDTD/extensions/engine_ext_mur_abc.cpp:28-178
 *   ems_add_lorentz       Engine_Ext_LorentzMaterial / Engine_Ext_Dispersive
 *                                                    FDTD/extensions/engine_ext_lorentzmaterial.cpp:79-173
 *   ems_add_lumped_rlc    Engine_Ext_LumpedRLC                            FDTD/extensions/engine_ext_lumpedRLC.cpp:83-147
 *   ems_iterate           Engine::IterateTS                              FDTD/engine.cpp:267-286
 *   ems_set_graph         Engine::IterateTS loop replayed as HIP graphs per timestep chunk
 *   ems_set_fused         Engine::IterateTS step form (single sweep or two half steps)
 *   ems_get_fused                                                        FDTD/engine.cpp:267-286
 *   ems_set_quad/get_quad Engine::UpdateVoltages / UpdateCurrents lane width (four x cells per lane)
 *                                                    FDTD/engine.cpp:110-222, FDTD/engine_sse.cpp:79-274
 *   ems_get_numts         Engine::GetNumberOfTimesteps                   FDTD/engine.h:74
 *   ems_get_volt/curr     Engine::GetVolt / GetCurr (bulk)                FDTD/engine.h:77-93
 *   ems_set_volt/curr     Engine::SetVolt / SetCurr (bulk)                FDTD/engine.h:95-101
 *   ems_read/write_field  Engine::GetVolt / GetCurr / SetVolt / SetCurr over a box
 *                                                                        FDTD/engine.h:77-101
 *   ems_get_field_at      Engine_Interface_FDTD::GetEField / GetHField   FDTD/engine_interface_fdtd.cpp:63-204
 *   ems_add_probe         ProcessVoltage / ProcessCurrent / ProcessFieldProbe sampling
 *                                                    Common/processvoltage.cpp:37-47, Common/processcurrent.cpp:96-171,
 *                                                    Common/processfieldprobe.cpp:77-92
 *   ems_add_fd_dump       ProcessFieldsFD::Process                       Common/processfields_fd.cpp:72-107
 *   ems_add_td_dump       ProcessFieldsTD::Process                       Common/processfields_td.cpp:50-91
 *   ems_energy            Engine_Interface_FDTD::CalcFastEnergy          FDTD/engine_interface_fdtd.cpp:302-347
 *   ems_energy_noalloc    (no counterpart: ems_energy for callers that cannot allocate)
 *   ems_set_steady_state  Operator_Ext_SteadyState -> Engine_Ext_SteadyState
 *   ems_steady_state                                 FDTD/extensions/engine_ext_steadystate.cpp:50-107
 *   ems_destroy           Engine::~Engine / Engine::Reset                FDTD/engine.cpp:45-108
 *   ems_adj_*             no counterpart: discrete adjoint of Engine::IterateTS for gradients
 *   ems_batch_*           no counterpart: many independent Engine::IterateTS loops
 *                         (openems.cpp:1388-1494, one per instance) advanced by one
 *                         launch per half step
 *
 * Conventions: every int-returning function returns EMS_OK (0) or a negative
 * EMS_E_* code and never exits the process; ems_last_error() gives the
 * message for the calling thread.  A handle is single-caller and not
 * re-entrant.  Work is queued on the handle's HIP stream; ems_iterate returns
 * after the queued steps have completed unless ems_set_async(h, 1) was called.
 */
#ifndef EMS_HIP_H
#define EMS_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define EMS_OK 0
#define EMS_E_ARG (-1)
#define EMS_E_STATE (-2)
#define EMS_E_HIP (-3)
#define EMS_E_NOMEM (-4)
#define EMS_E_NODEV (-5)

typedef struct ems_engine ems_engine;

const char* ems_last_error(void);
/* number of visible HIP devices (0 when no GPU is present) */
int ems_device_count(void);

/* Device coefficient layouts.
 *   PLAIN      : vv, vi, ii, iv as float32 arrays (24 B per cell per half step).
 *   COMPRESSED : one uint16 per cell into a table of unique coefficient tuples
 *                (the per-cell form of the reference's compressed operator;
 *                2 B per cell per half step); UPML coefficients likewise by a
 *                per-position table index.  Possible when the operator has at
 *                most 65536 unique tuples; results are bit-identical to PLAIN.
 *   AUTO       : COMPRESSED when possible and the table is reused on average
 *                by at least 16 cells, otherwise PLAIN (UPML decided alike). */
#define EMS_LAYOUT_AUTO 0
#define EMS_LAYOUT_PLAIN 1
#define EMS_LAYOUT_COMPRESSED 2

/* numLines[3]; vv, vi, ii, iv: host float32 [3][Nx][Ny][Nz] (reference N-I-J-K).
 * device: HIP device ordinal.  Uses EMS_LAYOUT_AUTO. */
int ems_create(ems_engine** out, const unsigned numLines[3], const float* vv, const float* vi,
               const float* ii, const float* iv, int device);

/* As ems_create with an explicit layout; EMS_LAYOUT_COMPRESSED fails with
 * EMS_E_ARG when the operator has more than 65536 unique tuples. */
int ems_create_layout(ems_engine** out, const unsigned numLines[3], const float* vv, const float* vi,
                      const float* ii, const float* iv, int device, int layout);

/* From the reference's SSE-compressed operator: op_index is m_Op_index
 * ([Nx][Ny][Nz] unsigned, only z < nV = ceil(Nz/4) used), n_unique its table
 * size and vv4[n], vi4[n], ii4[n], iv4[n] (n = 0..2) the f4_*_Compressed[n]
 * tables (n_unique x 4 floats, lane l of entry u holds z = l*nV + v). */
int ems_create_sse_compressed(ems_engine** out, const unsigned numLines[3], const unsigned* op_index,
                              unsigned n_unique, const float* const vv4[3], const float* const vi4[3],
                              const float* const ii4[3], const float* const iv4[3], int device, int layout);

/* Host only (no device needed): the per-cell compression ems_create applies.
 * idx receives Nx*Ny*Nz uint16 in x-fastest order (x + Nx*(y + Ny*z));
 * tabV / tabI receive 8 floats per unique tuple, {vv_x, vi_x, vv_y, vi_y,
 * vv_z, vi_z, 0, 0} and {ii_x, iv_x, ii_y, iv_y, ii_z, iv_z, 0, 0}, numbered
 * in first-seen order over cells x -> y -> z (x outermost), tuples equal by
 * float value.  Returns the number of unique tuples, or EMS_E_ARG when there
 * are more than max_unique (at most 65536). */
int ems_compress_operator(const unsigned numLines[3], const float* vv, const float* vi, const float* ii,
                          const float* iv, unsigned short* idx, float* tabV, float* tabI, unsigned max_unique);

/* Host only: the N-I-J-K operator arrays [3][Nx][Ny][Nz] of the reference's
 * SSE-compressed operator (arguments as in ems_create_sse_compressed). */
int ems_expand_sse_compressed(const unsigned numLines[3], const unsigned* op_index, unsigned n_unique,
                              const float* const vv4[3], const float* const vi4[3], const float* const ii4[3],
                              const float* const iv4[3], float* vv, float* vi, float* ii, float* iv);

/* Steady-state detection (Operator_Ext_SteadyState / Engine_Ext_SteadyState,
 * FDTD/extensions/engine_ext_steadystate.cpp:50-107, set up by
 * openems.cpp:1205-1234 for periodic excitations), before ems_init.
 * period_steps: m_TS_period = round(period / dT); probes dir[n], pos[3][n]
 * (E probes on voltage edges).  Every step records the probe voltages after
 * the voltage update (before the other apply hooks); every period from 2p on
 * it sets the last max diff: the largest relative power change between the
 * last two periods over the probes holding at least 1 % of the largest
 * power, and the relative energy change.  history_capacity: evaluations kept. */
int ems_set_steady_state(ems_engine* h, unsigned period_steps, unsigned nprobe, const int* dir, const unsigned* pos,
                         unsigned history_capacity);
/* Engine_Ext_SteadyState::GetLastDiff and the evaluation history (up to cap). */
int ems_steady_state(ems_engine* h, double* last_max_diff, unsigned* evaluations, unsigned* ts, double* diff,
                     unsigned cap);

/* Step form, before ems_init.  A single-sweep form advances a whole timestep
 * in one launch over double-buffered fields (voltage update, voltage sources,
 * current update, current sources, UPML in both halves), bit-identical to the
 * two-half-step form with about 40 % fewer algorithmic bytes.  Form 1 stages
 * V^{n+1} planes of 64 x 16 column tiles in LDS; form 2 sweeps independent
 * columns, recomputing the neighbour voltages each current update needs.
 * Both need every active extension to have a single-sweep form: Mur,
 * Lorentz / Drude, lumped RLC and steady-state detection keep two half steps,
 * as do current sources on the last line of a direction and more than 64
 * sources in one tile (form 1) or column (form 2).
 * Form 3 sweeps columns of four x cells per lane (16-byte accesses) like
 * form 2, on meshes without UPML whose x lines are a multiple of 4, Mur
 * faces included; on rows of at least 64 quads it runs in 64 x 4 quad tiles
 * that take neighbour voltages from LDS and adjacent lanes, and runs the
 * pre and post hooks of Mur faces on x = 0 and x = Nx-1 itself.
 * mode: -1 automatic (form 3 on meshes without UPML whose rows hold at
 * least 64 quads, where every active extension has a sweep form and the
 * second field copy fits in device memory; two half steps otherwise),
 * 0 two half steps, 1 LDS sweep, 2 column sweep, 3 quad column sweep
 * (ems_init then fails with EMS_E_ARG when the engine has no such form). */
int ems_set_fused(ems_engine* h, int mode);
/* fused: active form (0 two half steps, 1 LDS sweep, 2 column sweep, 3 quad sweep);
 * zdepth: z planes per workgroup */
int ems_get_fused(const ems_engine* h, int* fused, unsigned* zdepth);

/* Half-step lane width (Engine::UpdateVoltages / UpdateCurrents,
 * FDTD/engine.cpp:110-222; the reference's SSE engine packs four z cells per
 * vector, FDTD/engine_sse.cpp:79-274).  With quads on, each lane of the two
 * half steps updates four consecutive x cells with 16-byte accesses; results
 * are bit-identical to one cell per lane.  Available when the number of x
 * lines is a multiple of 4 and the engine runs two half steps.
 * mode: -1 automatic (the default: quads where available on meshes without
 * UPML), 0 one cell per lane, 1 quads (ems_init fails with EMS_E_ARG where
 * unavailable). */
int ems_set_quad(ems_engine* h, int mode);
/* on: 1 when the half steps run quads; zdepth: their z planes per workgroup */
int ems_get_quad(const ems_engine* h, int* on, unsigned* zdepth);

/* Active layouts and table sizes (known after ems_init for the UPML part):
 * layout, pml_layout: EMS_LAYOUT_PLAIN or EMS_LAYOUT_COMPRESSED (pml_layout 0
 * without UPML); n_unique / n_unique_pml: table entries (0 when plain). */
int ems_get_layout(const ems_engine* h, int* layout, unsigned* n_unique, int* pml_layout,
                   unsigned* n_unique_pml);

/* Excitation signal (length samples, volt at n*dT, curr at (n+1/2)*dT),
 * period in timesteps (0 when aperiodic) and the voltage / current cell lists:
 * idx [3][K] (x row, y row, z row), dir [K], amp [K], delay [K] (timesteps). */
int ems_set_excitation(ems_engine* h, const float* sig_volt, const float* sig_curr, unsigned length,
                       int period_steps,
                       unsigned nvolt, const unsigned* vidx, const unsigned short* vdir, const float* vamp,
                       const unsigned* vdelay,
                       unsigned ncurr, const unsigned* cidx, const unsigned short* cdir, const float* camp,
                       const unsigned* cdelay);

/* One UPML slab in creation order: start[3], num[3] (lines), six arrays
 * [3][num0][num1][num2]: vv, vvfo, vvfn, ii, iifo, iifn. */
int ems_add_upml(ems_engine* h, const unsigned start[3], const unsigned num[3], const float* vv,
                 const float* vvfo, const float* vvfn, const float* ii, const float* iifo,
                 const float* iifn);

/* One Mur face in creation order: normal ny, face line, neighbour line,
 * coefficients [N[ny+1]][N[ny+2]] for the two tangential components, and the
 * first timestep at which the boundary is active. */
int ems_add_mur(ems_engine* h, int ny, unsigned line, unsigned line_shift, const float* coeff_nyP,
                const float* coeff_nyPP, unsigned start_ts);

/* Lumped RLC elements (ADE form): pos [3][count], dir [count], and nine
 * coefficient arrays [count]: i2v, ilv, vvd, vv2, vj1, vj2, ib0, b1, b2. */
int ems_add_lumped_rlc(ems_engine* h, unsigned count, const unsigned* pos, const int* dir,
                       const float* const coeffs[9]);

/* Lorentz/Drude/Debye ADE of one dispersion order: cell list pos [3][count],
 * coefficients [3][count] (v_int, v_ext, v_lor, i_int, i_ext, i_lor; the
 * *_lor arrays may be NULL when the order has no Lorentz pole). */
int ems_add_lorentz(ems_engine* h, unsigned count, const unsigned* pos, const float* v_int,
                    const float* v_ext, const float* v_lor, const float* i_int, const float* i_ext,
                    const float* i_lor, int volt_on, int curr_on);

/* Finish set-up (extension ordering, device tables).  Called implicitly by
 * the first ems_iterate. */
int ems_init(ems_engine* h);

int ems_iterate(ems_engine* h, unsigned nsteps);
unsigned ems_get_numts(const ems_engine* h);
int ems_set_async(ems_engine* h, int async);

/* Timestep chunks as HIP graphs (the IterateTS loop, FDTD/engine.cpp:267-286,
 * with its ProcessingArray sampling).  ems_iterate replays every full chunk of
 * chunk_steps steps from a graph captured from the direct launches the first
 * time that chunk's launch pattern occurs (probe, dump and steady-state
 * decisions of its steps, buffer parity of the single-sweep forms).  The
 * chunk is rounded up to a multiple of the pattern's period (the sampling
 * intervals, the steady-state period) when that period is at most 1024
 * steps, so chunks repeat; at most 16 patterns are kept, further ones launch
 * directly.  Remaining steps and profiled steps launch directly.  Default 32
 * steps; 0 launches every step directly.  Results are identical either way.
 * Adding probes or dumps drops the captured graphs. */
int ems_set_graph(ems_engine* h, unsigned chunk_steps);
/* chunk length, graphs captured and chunks replayed so far */
int ems_get_graph(const ems_engine* h, unsigned* chunk_steps, unsigned* captures, unsigned* replays);
int ems_sync(ems_engine* h);

/* whole-field transfers in the reference N-I-J-K layout, [3][Nx][Ny][Nz] */
int ems_get_volt(ems_engine* h, float* out);
int ems_get_curr(ems_engine* h, float* out);
int ems_set_volt(ems_engine* h, const float* in);
int ems_set_curr(ems_engine* h, const float* in);

/* Engine::GetVolt / GetCurr and SetVolt / SetCurr (FDTD/engine.h:77-101) over
 * the box start..stop (inclusive mesh indices): comp 0, 1 or 2 for one
 * component, data [nx][ny][nz]; comp -1 for all three, data [3][nx][ny][nz]
 * (N-I-J-K, z fastest).  Only the box travels over PCIe.  Writes reach both
 * buffers of a single-sweep engine. */
int ems_read_field(ems_engine* h, int is_curr, int comp, const unsigned start[3], const unsigned stop[3], float* out);
int ems_write_field(ems_engine* h, int is_curr, int comp, const unsigned start[3], const unsigned stop[3],
                    const float* in);
/* Engine_Interface_FDTD::GetEField / GetHField (FDTD/engine_interface_fdtd.cpp:63-204,
 * GetRawField :263-268): field 0 (E, from the voltages) or 1 (H, from the
 * currents) at mesh node pos, interp 0 none, 1 node, 2 cell interpolation
 * (Engine_Interface_Base::InterpolationType); edge_len[3]: the edge lengths
 * per axis in metres (primary lines for E, dual for H); out[3] in V/m or A/m.
 * The same formulas as the field dumps (ems_add_fd_dump / ems_add_td_dump). */
int ems_get_field_at(ems_engine* h, int field, int interp, const unsigned pos[3], const double* const edge_len[3],
                     double out[3]);

/* ---- probes (evaluated on the device at their sampling steps) ---- */
#define EMS_PROBE_VOLTAGE 0   /* line integral of voltages, double accumulation   */
#define EMS_PROBE_CURRENT 1   /* closed loop of currents, float accumulation       */
#define EMS_PROBE_EFIELD 2    /* E = V / edge length at one node (3 components)    */
#define EMS_PROBE_HFIELD 3    /* H = I / dual edge length at one node              */

typedef struct ems_probe_desc {
    int type;                     /* EMS_PROBE_*                                      */
    unsigned nelem;               /* voltage/current: number of terms; points: 3       */
    const unsigned* comp;         /* [nelem] field component of each term               */
    const unsigned* pos;          /* [3][nelem] mesh index of each term                 */
    const int* sign;              /* [nelem] +1 / -1 (voltage, current)                 */
    const double* divisor;        /* [nelem] edge lengths (points), else NULL           */
    unsigned interval;            /* time-domain sampling interval, 0 = none            */
    unsigned fd_interval;         /* DFT sampling interval, 0 = none                    */
    unsigned start_ts, stop_ts;   /* active window (Processing::SetProcessStartStopTime)*/
    unsigned capacity;            /* maximum number of samples to record                */
} ems_probe_desc;

int ems_add_probe(ems_engine* h, const ems_probe_desc* d, int* id);
/* number of recorded samples; with ts/values non-NULL copies them out:
 * ts[count], values[count*3] (slot 0 used by voltage/current probes) */
int ems_probe_results(ems_engine* h, int id, unsigned* count, unsigned* ts, double* values);

/* ---- frequency-domain field dump (ProcessFieldsFD, dump types 10 and 11) ----
 * field 0 = E (voltage / edge length), 1 = H (current / dual edge length);
 * interp 0 = none, 1 = node, 2 = cell interpolation
 * (Engine_Interface_FDTD::GetRawInterpolatedField / ...DualField,
 * FDTD/engine_interface_fdtd.cpp:63-204).  Box start/stop are mesh indices
 * (snapped by the caller, on the dual mesh where the reference uses it);
 * edge_len[3] are the primary (E) or dual (H) edge lengths per axis.  The
 * nfreq complex<float> factors per sample are supplied up front:
 * factors[nsamples][nfreq][2] for samples taken at ts = first_ts + k*fd_interval
 * (k = 0..nsamples-1; first_ts a multiple of fd_interval, the first sampling
 * step inside the dump's start/stop window), with the half-step time shift
 * of H already in them. */
int ems_add_fd_dump(ems_engine* h, int field, int interp, const unsigned start[3], const unsigned stop[3],
                    unsigned fd_interval, unsigned first_ts, unsigned nfreq, unsigned nsamples,
                    const float* factors, const double* const edge_len[3], int* id);
/* out: [nfreq][3][nx][ny][nz][2] float32 (re, im), dump-box N-I-J-K order */
int ems_fd_dump_results(ems_engine* h, int id, float* out);

/* ---- time-domain field dump (ProcessFieldsTD::Process, Common/processfields_td.cpp:50-91;
 * dump types 0 and 1) ----
 * The E (field 0) or H (field 1) field of the box start..stop, interpolated as
 * for ems_add_fd_dump, stored after every step ts with ts % interval == 0 and
 * start_ts <= ts <= stop_ts (Processing::CheckTimestep), at most capacity
 * samples, all kept on the device until read. */
int ems_add_td_dump(ems_engine* h, int field, int interp, const unsigned start[3], const unsigned stop[3],
                    unsigned interval, unsigned start_ts, unsigned stop_ts, unsigned capacity,
                    const double* const edge_len[3], int* id);
/* *count = samples taken so far; ts[count] their steps; out[count][3][nx][ny][nz]
 * float32 (dump-box N-I-J-K order); ts and out may be NULL */
int ems_td_dump_results(ems_engine* h, int id, unsigned* count, unsigned* ts, float* out);

/* eps0 * sum(V^2) + mu0 * sum(I^2) over [0,N-1)^3 */
int ems_energy(ems_engine* h, double* out);

/* ems_energy without allocation: no counterpart in the reference.  For callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Enqueues the same
 * reduction on the engine's stream and writes the same value (bit for bit)
 * to *d_out, one double in device memory; it neither allocates nor
 * synchronises: the value is ready once the engine's stream reaches it
 * (ems_sync).  The partial sums live in a file-scope device array of 2048
 * doubles (16 KiB, one reduction workgroup's pair each; the workgroup count
 * is fixed, so every mesh size fits and none is refused), held by each
 * device until the process ends and shared by every call on that device:
 * calls that overlap on one GPU (engines on different streams) race. */
int ems_energy_noalloc(ems_engine* h, double* d_out);

/* device time of the last ems_iterate call in milliseconds (HIP events) and
 * the average duration of the voltage / current update kernels it launched */
int ems_last_timing(ems_engine* h, double* total_ms, double* volt_ms, double* curr_ms);

/* record HIP events around every voltage / current update kernel (adds one
 * host synchronisation per step; used for kernel roofline measurements) */
int ems_set_profiling(ems_engine* h, int enable);

void ems_destroy(ems_engine* h);

/* ---- batched engine --------------------------------------------------------
 * Many independent engines (own mesh, operator, UPML, sources, probes) advance
 * together: each half step is one persistent launch over a work table of
 * (instance, tile) items, plus one launch each for the sources and the probes
 * of all instances.  Every instance stays bit-identical to running its engine
 * alone.  The engines must be initialised, share a device, a coefficient
 * layout and a UPML layout, run two half steps (ems_set_fused(h, 0)), stand at
 * the same timestep, and carry no Mur, Lorentz, lumped RLC or field dump.  The
 * batch does not own them: destroy it before them.  While batched, advance the
 * engines only through ems_batch_iterate; ems_probe_results / ems_get_volt /
 * ems_get_curr on each engine read its results. */
typedef struct ems_batch ems_batch;
int ems_batch_create(ems_batch** out, ems_engine* const* engines, unsigned count);
/* tile_x: 0 = 256 consecutive plane positions per workgroup, else tile_x x
 * (256/tile_x) columns; zdepth[count]: z planes per workgroup and instance
 * (NULL: each instance's default); grid: workgroups per launch (0: one per
 * work item, else a persistent grid that strides over the items). */
int ems_batch_set_config(ems_batch* b, unsigned tile_x, const unsigned* zdepth, unsigned grid);
int ems_batch_get_config(const ems_batch* b, unsigned* tile_x, unsigned* zdepth, unsigned* grid, unsigned* nitems);
/* Instances are stepped in groups: every group advances all steps of an
 * ems_batch_iterate call before the next group starts, so a group whose
 * device working set fits the last-level cache stays resident across steps.
 * group_bytes: working-set budget of a group, packed first-fit by decreasing
 * size (0: all instances in one group, the default). */
int ems_batch_set_grouping(ems_batch* b, unsigned long long group_bytes);
int ems_batch_get_grouping(const ems_batch* b, unsigned long long* group_bytes, unsigned* ngroups);
/* Time every candidate configuration (tile shape, z depths, grid, group
 * budget) for `steps` steps of every group on the batch itself, restore the
 * advanced state, and keep the fastest (best_ms: its time per step). */
int ems_batch_autotune(ems_batch* b, unsigned steps, double* best_ms);
int ems_batch_iterate(ems_batch* b, unsigned nsteps);
/* Each group's launches of `steps` steps are captured once as a HIP graph
 * and replayed (0: direct launches, the default).  Results are the same. */
int ems_batch_set_graph(ems_batch* b, unsigned steps);
int ems_batch_get_graph(const ems_batch* b, unsigned* steps);
/* Run the groups on n streams (1..4; group i on stream i % n), so that
 * several groups advance at once; each group's working set should then be
 * about the cache budget over n.  ems_batch_iterate returns when all are done. */
int ems_batch_set_streams(ems_batch* b, unsigned n);
int ems_batch_get_streams(const ems_batch* b, unsigned* n);
/* device time of the last ems_batch_iterate call */
int ems_batch_last_ms(const ems_batch* b, double* ms);
void ems_batch_destroy(ems_batch* b);

/* ---- adjoint gradients ------------------------------------------------------
 * No counterpart in the reference: the discrete adjoint of the engine's
 * timestep (two half steps with UPML and sources) for the objective
 *   J = sum_p |F_p|^2,  F_p = sum_{n=1..N} (V[comp_p](pos_p) / divisor_p)(n) exp(-2 pi i f n dT) dT
 * (V after step n, N steps from the engine's state at ems_adj_forward).
 * ems_adj_forward runs the N steps, saving the state every checkpoint_every
 * steps; ems_adj_backward replays each segment from its checkpoint and runs
 * the transposed half steps backwards in time, accumulating dJ/dvi and dJ/dvv
 * (double, [3][box]) on the edges of the inclusive sensitivity box
 * box_lo..box_hi, which must not overlap the UPML.  The engine runs two half
 * steps and carries no Mur, Lorentz or lumped RLC; its probes and dumps are
 * not sampled during these runs.  Afterwards the engine holds the state at
 * the end of the forward run. */
typedef struct ems_adjoint ems_adjoint;
int ems_adj_create(ems_adjoint** out, ems_engine* h, unsigned nprobe, const unsigned* comp, const unsigned* pos,
                   const double* divisor, double freq, double dT, const unsigned box_lo[3], const unsigned box_hi[3],
                   unsigned checkpoint_every);
int ems_adj_forward(ems_adjoint* a, unsigned nsteps, double* J);
int ems_adj_backward(ems_adjoint* a, double* dJ_dvi, double* dJ_dvv);
void ems_adj_destroy(ems_adjoint* a);

#ifdef __cplusplus
}
#endif
#endif /* EMS_HIP_H */
