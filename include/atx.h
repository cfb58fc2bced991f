/*
 * atx.h -- C ABI of the MI355X tracking engine (libatx.so).
 *
 * The engine replaces the GPU tracking backend of AT: the Python extension
 * at.tracking.gpu (atgpu/PyInterface.cpp:110-402, stub
 * pyat/at/tracking/gpu.pyi:9-28) that pyat's lattice_track/internal_lpass
 * call when use_mp=MPMode.GPU (pyat/at/tracking/track.py:115-127). It keeps
 * the reference plugin surface: pass methods are named as in
 * PassMethodFactory (atgpu/PassMethodFactory.cpp:33-46) and their
 * parameters are given by attribute name, with the names, required/optional
 * status and defaults of the CPU integrators (atintegrators/<Pass>.c
 * trackFunction getters).
 *
 * All entry points return ATX_OK (0) or an error code; atx_last_error()
 * returns the message of the last failing call on the calling thread. The
 * codes map onto the Python exceptions pyat raises for the same conditions
 * (pyat/at.c:409-416,491-494,514-536; atgpu/PyInterface.cpp:292-297).
 *
 * Arrays use the reference layouts: particle coordinates are (6, N)
 * Fortran-ordered doubles (coordinate index fastest, pyat/at.c:409-418), refpt
 * output is (6, N, R, T) Fortran-ordered (at.c:567-572), matrices R1/R2 are
 * 6x6 column-major (atlalib.c:17-32).
 */
#ifndef ATX_H
#define ATX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ATX_OK 0
#define ATX_E_VALUE 1     /* ValueError: shape, range, partial-turn or energy errors */
#define ATX_E_ATTRIBUTE 2 /* AttributeError: a required element attribute is missing */
#define ATX_E_RUNTIME 3   /* RuntimeError: unknown PassMethod, device or launch failure */
#define ATX_E_NODEVICE 4  /* no HIP device available */

/* atx_track flags */
#define ATX_DEVICE_POINTERS 0x1u /* all arrays are device pointers; run asynchronously on stream */
#define ATX_STRICT 0x2u          /* reference operation order, no FMA contraction */
#define ATX_KEEP_COUNTER 0x4u    /* continue from the turn counter saved by the previous call */
#define ATX_SOA 0x8u             /* with ATX_DEVICE_POINTERS: r_in is SoA (x[N], px[N], ...) */
#define ATX_CONTINUE_LOSSES 0x10u /* atx_track_batch: the loss arrays hold the map of earlier
                                     calls on the same particles; they are not reset, and
                                     particles already flagged are not recorded again */
#define ATX_REPACK 0x20u         /* atx_track_batch, and atx_track without refpts: track in
                                  * turn blocks of 1, 1, 2, 4, ... turns and, between blocks,
                                  * pack each instance's surviving particles into its leading
                                  * slots and the workgroups that hold them, over all
                                  * instances, at the front of the grid; the blocks run as one
                                  * hipGraph. Results are identical to a plain launch. */

typedef struct atx_lattice atx_lattice;

/* Message of the last failing call on this thread ("" if none). */
const char *atx_last_error(void);

/* --- Lattice construction ------------------------------------------------
 * Replaces Lattice::addElement -> PassMethodFactory::createElement ->
 * <Pass>::getParameters(AbstractInterface*) (atgpu/Lattice.cpp:37-56,
 * atgpu/PassMethodFactory.cpp:73-94, atgpu/AbstractInterface.h:9-43).
 * Attributes are set on the element added last. An unknown pass method is
 * rejected when added ("Not implemented PassMethod: X", Lattice.cpp:45-54);
 * missing required attributes are reported by atx_track/atx_lattice_check.
 */
int atx_lattice_create(atx_lattice **out);
int atx_lattice_destroy(atx_lattice *lat);
int atx_lattice_add_element(atx_lattice *lat, const char *pass_method, const char *fam_name);
int atx_element_set_double(atx_lattice *lat, const char *name, double value);
int atx_element_set_int(atx_lattice *lat, const char *name, int64_t value);
int atx_element_set_array(atx_lattice *lat, const char *name, const double *data, int64_t n);
int atx_lattice_num_elements(const atx_lattice *lat);
/* Symplectic integrator of the thick multipole slices, the `integrator`
 * keyword of gpupass (atgpu/PyInterface.cpp:44-50,259-266) with the
 * coefficient sets of atgpu/SymplecticIntegrator.cpp:32-86: 1 Euler,
 * 2 McLachlan-Atela 2nd order, 3 Ruth 3rd order, 4 Forest-Ruth 4th order
 * (default; the only one the CPU integrators use), 5 McLachlan-Atela 4th
 * order, 6 Yoshida 6th order. Other values: ATX_E_RUNTIME. */
int atx_lattice_set_integrator(atx_lattice *lat, int32_t type);

/* Validate every element against its pass method (required attributes,
 * array sizes, energy where radiation or RF needs it) without tracking. */
int atx_lattice_check(atx_lattice *lat, double energy, double rest_energy);

/* --- Tracking ------------------------------------------------------------
 * Replaces at_gpupass (atgpu/PyInterface.cpp:132-402) with the loss and turn
 * semantics of the CPU driver at_atpass (pyat/at.c:340-681):
 *   r_in       (6,N) F-order, updated in place with the final coordinates;
 *   refpts     sorted uint32 element indices (0 = entrance of the first
 *              element, nelem = exit of the last), copies written to r_out
 *              (6,N,R,T) in the order at.c writes them; r_out may be NULL
 *              when num_refpts == 0;
 *   turn       starting value of the turn counter (ignored with
 *              ATX_KEEP_COUNTER); the counter enters RFCavityPass and the
 *              loss turn;
 *   energy     ring energy [eV]; rest_energy [eV] (0: relativistic);
 *   losses     non-zero: fill lost_flag/lost_turn/lost_elem/lost_coord
 *              (N, N, N, (6,N)); survivors keep 0 (at.c:574-594);
 *   start_elem/end_elem  partial-turn tracking (end_elem 0 = all);
 *   stream     hipStream_t (NULL = default stream);
 *   flags      ATX_* flags above.
 * Without ATX_DEVICE_POINTERS the arrays are host memory and the call
 * returns when the results are back on the host. charge is accepted for
 * interface parity and unused by the in-scope pass methods. */
int atx_track(atx_lattice *lat, double *r_in, uint32_t num_particles, int32_t nturns,
              const uint32_t *refpts, uint32_t num_refpts, int32_t turn, double energy,
              double rest_energy, double charge, int32_t losses, int32_t start_elem,
              int32_t end_elem, double *r_out, uint8_t *lost_flag, uint32_t *lost_turn,
              uint32_t *lost_elem, double *lost_coord, void *stream, uint32_t flags);

/* Single-element pass of element `index` with the semantics of at_elempass
 * (pyat/at.c:683-753): turn 0, ring length 0, no loss check. */
int atx_elempass(atx_lattice *lat, int32_t index, double *r_in, uint32_t num_particles,
                 double energy, double rest_energy, void *stream, uint32_t flags);

/* Launch shape of the tracking kernel (no counterpart in the reference,
 * whose CUDA launch is fixed at 64 threads per block, one particle per
 * thread: atgpu/CudaGPU.cpp:157-185). block: lanes per workgroup, a
 * multiple of 64 in [64, 512]; ppt: particles per lane, 1 or 2 (a lane
 * steps its particles through each magnet together); unroll: slices per
 * loop iteration of the thick multipoles, 1, 2 or 4 (1 or 2 with ppt 2). A
 * zero field takes the engine's default. Results do not depend on the
 * shape. */
typedef struct atx_launch_shape {
    int32_t block;
    int32_t ppt;
    int32_t unroll;
    int32_t pad;
} atx_launch_shape;

/* Shape of the atx_track / atx_elempass launches on this lattice (NULL:
 * back to the engine's default). */
int atx_lattice_set_launch_shape(atx_lattice *lat, const atx_launch_shape *shape);

/* --- Batched instances ---------------------------------------------------
 * Many lattice instances (e.g. error realisations), each with its own
 * device-resident particles, tracked in ONE launch: workgroups are assigned
 * to instances in proportion to their particle counts, and the instances
 * with the most turns x elements per workgroup are dispatched first. Each
 * item tracks the call's nturns from the call's turn, or its own
 * (nturns, turn) when item.nturns >= 0. No counterpart in the reference,
 * which tracks one lattice per call (at.c:340, PyInterface.cpp:132); each
 * instance has the semantics of atx_track with refpts = [] and start/end =
 * whole ring. All pointers are device pointers; the call is asynchronous on
 * stream. shape: launch shape (NULL: the engine's default). */
typedef struct atx_batch_item {
    atx_lattice *lattice;
    double *r;              /* (6,N) F-order, or SoA with ATX_SOA */
    uint32_t num_particles;
    uint32_t pad;
    uint8_t *lost_flag;     /* N, or NULL when losses == 0 */
    uint32_t *lost_turn;    /* N */
    uint32_t *lost_elem;    /* N */
    double *lost_coord;     /* (6,N) */
    int32_t nturns;         /* turns of this item; < 0: the call's nturns and turn */
    int32_t turn;           /* first turn number of this item when nturns >= 0 */
} atx_batch_item;

int atx_track_batch(const atx_batch_item *items, int32_t num_items, int32_t nturns, int32_t turn,
                    double energy, double rest_energy, int32_t losses,
                    const atx_launch_shape *shape, void *stream, uint32_t flags);

/* atx_track_batch without any allocation. No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). Same
 * arguments and results as atx_track_batch, with these limits: at most 64
 * items; no ATX_REPACK (its survivor lists need working memory); every
 * item's lattice already resident on the current device for this energy
 * and rest energy (track it once with atx_track or atx_track_batch first).
 * The instance table lives in a fixed device array of the library, one per
 * GPU (64 entries, about 10 KB), written by each call: calls that overlap on
 * one GPU (on different streams, or from several threads) race on it. Calls
 * no hipMalloc, hipFree or other allocating routine. */
int atx_track_batch_noalloc(const atx_batch_item *items, int32_t num_items, int32_t nturns,
                            int32_t turn, double energy, double rest_energy, int32_t losses,
                            const atx_launch_shape *shape, void *stream, uint32_t flags);

/* --- Differentiation -----------------------------------------------------
 * Reverse-mode gradient through nturns turns of the whole ring. No
 * counterpart in the reference, whose tangent maps (diff_drift.c,
 * diff_thinkick.c, diff_bnd_kick.c, diff_bend_fringe.c) only propagate the
 * diffusion matrix (at.c:755-831); this entry applies the transposes of the
 * same per-step Jacobians.
 *   states   device, (nturns+1) consecutive (6,N) F-order blocks: the state
 *            at the start of each turn (as atx_track writes them with
 *            refpts = [0]) followed by the state after the last turn;
 *   lam      device (6,N): dL/d(end state) in, dL/d(start state) out;
 *            lanes whose end state is lost are skipped;
 *   grad_slot/grad_coeff  host arrays of nelem ints: element e contributes
 *            dL/dPolynomB[grad_coeff[e]] to grad[grad_slot[e]] (slot -1:
 *            not differentiated);
 *   grad     device array of num_params doubles, accumulated (+=).
 * Supported: IdentityPass, DriftPass, AperturePass, CorrectorPass,
 * RFCavityPass, ThinMPolePass and the non-radiating Str/BndMPole passes
 * with integrator 4, FieldScaling 1, no quadrupole fringe, edge fringe
 * without gap correction, and misalignments that leave delta unchanged;
 * anything else is ATX_E_RUNTIME. Asynchronous on stream. */
int atx_adjoint(atx_lattice *lat, const double *states, int32_t nturns, int32_t turn,
                uint32_t num_particles, double energy, double rest_energy, double *lam,
                const int32_t *grad_slot, const int32_t *grad_coeff, double *grad,
                int32_t num_params, void *stream);

/* Several independent blocks of the same lattice in one launch (no
 * counterpart in the reference). Each item is one atx_adjoint call: its
 * own turn-start states, adjoint (in place), first turn and particle
 * count; `accumulate` nonzero adds its dL/dPolynomB to grad, zero only
 * propagates lam (e.g. unit seeds, giving rows of the block's transposed
 * Jacobian). grad may be NULL when no item accumulates or no element has a
 * gradient slot. At most 65,535 items. Asynchronous on stream. */
typedef struct atx_adjoint_item {
  const double *states; /* device, (nturns+1) (6,N) blocks */
  double *lam;          /* device (6,N), in place */
  int32_t nturns;
  int32_t turn;
  uint32_t num_particles;
  int32_t accumulate;
} atx_adjoint_item;

int atx_adjoint_batch(atx_lattice *lat, const atx_adjoint_item *items, int32_t nitems,
                      double energy, double rest_energy, const int32_t *grad_slot,
                      const int32_t *grad_coeff, double *grad, int32_t num_params,
                      void *stream);

/* --- Diffusion matrices --------------------------------------------------
 * Replaces at_diffmatrix (pyat/at.c:755-831), called once per element by
 * ohmi_envelope (pyat/at/physics/radiation.py:78-121): for every element of
 * the lattice at once, the Ohmi diffusion matrix B accumulated through the
 * element by its radiating pass from the orbit at its entrance, with the
 * reference's per-step propagation B <- M B M^T + D (diff_drift.c,
 * diff_thinkick.c, diff_str_kick.c, diff_bnd_kick.c, diff_bend_fringe.c,
 * diff_exactdrift.c, diff_str_exactkick.c). Elements without a radiating
 * pass (StrMPole/BndMPole/ExactMultipole Rad) get zeros.
 *   orbits  (6, nelem) F-order: orbit at the entrance of each element
 *   bdiff   nelem blocks of 36 doubles, column-major 6x6 ([i][j] at i+6j)
 * Host arrays (synchronous) or, with ATX_DEVICE_POINTERS, device arrays
 * (asynchronous on stream). One lane per element, one launch. */
int atx_diffusion_matrices(atx_lattice *lat, const double *orbits, double *bdiff, double energy,
                           double rest_energy, void *stream, uint32_t flags);

/* --- Element matrices ----------------------------------------------------
 * Replaces find_elem_m66 (pyat/at/physics/matrix.py:171-202), which
 * _dmatr calls once per element (pyat/at/physics/radiation.py:78-90), each
 * call an at_elempass of 12 particles (pyat/at.c at_elempass): for every
 * element of the lattice at once, its 6x6 matrix by central differences
 * around the orbit at its entrance. Probe j+ / j- is the orbit with +step/2
 * / -step/2 added to coordinate j (matrix.py:195-199), passed through that
 * element alone (no loss check); m[i][j] = (probe j+ - probe j-)_i / step.
 * One tracking launch over all elements (one instance each).
 *   orbits  (6, nelem) F-order: orbit at the entrance of each element
 *   m66     nelem blocks of 36 doubles, column-major 6x6 ([i][j] at i+6j)
 *   step    the reference's XYStep (nonzero)
 * Host arrays (synchronous) or, with ATX_DEVICE_POINTERS, device arrays
 * (asynchronous on stream). ATX_STRICT: the strict build. */
int atx_element_matrices(atx_lattice *lat, const double *orbits, double step, double *m66,
                         double energy, double rest_energy, void *stream, uint32_t flags);

/* Replace an attribute of element `index` (e.g. PolynomB of a magnet whose
 * strength is a differentiation parameter); the compiled lattice is
 * rebuilt on the next call. No counterpart in the reference, which
 * re-reads the Python elements unless reuse is set (at.c:445-512). */
int atx_element_set_array_at(atx_lattice *lat, int32_t index, const char *name,
                             const double *data, int64_t n);

/* Turn counter after the last call (at.c:656 last_turn). */
int atx_lattice_turn_counter(const atx_lattice *lat, int64_t *turn);

/* Release the HIP objects the engine keeps between calls: executable graphs
 * of repacked launches not yet destroyed and their events (waited for
 * first), and the pinned staging buffers of the launch tables with their
 * events. Lattice handles are not touched. The engine stays usable: later
 * calls create what they need again. For a process's end, before the HIP
 * runtime unloads (the Python binding registers it with atexit). No
 * counterpart in the reference, whose GPU backend keeps its objects in
 * static caches until exit (atgpu/PyInterface.cpp:135-137). */
int atx_shutdown(void);

/* --- Devices -------------------------------------------------------------
 * Replaces at_gpuinfo (atgpu/PyInterface.cpp:110-130): one record per
 * device: name, architecture (gfx name), compute-unit count, platform. */
int atx_device_count(int *count);
int atx_device_info(int device, char *name, int name_len, char *arch, int arch_len,
                    int *compute_units, char *platform, int platform_len);
int atx_set_device(int device);
int atx_get_device(int *device);

/* --- Device memory, pinned host memory, streams and events ----------------
 * Replaces the memory half of atgpu's GPU context (allocDevice, freeDevice,
 * hostToDevice, deviceToHost: atgpu/AbstractGPU.h:31-40, used by
 * Lattice::fillGPUMemory, atgpu/Lattice.cpp:272-303), and adds the stream
 * and event handles that asynchronous callers (the batching and serving
 * loops) need to keep state resident on the device and overlap copies with
 * tracking without a second runtime. Handles are opaque: `stream` and
 * `event` are a hipStream_t and a hipEvent_t; a NULL stream is the null
 * stream. Streams are created non-blocking, at the default priority.
 * atx_memcpy copies in either direction (the pointers say which) and is
 * asynchronous on `stream` (synchronous with respect to the host only for
 * pageable host memory). atx_event_query sets *done to 1 when the work
 * before the event has completed, 0 otherwise. */
int atx_device_alloc(void **ptr, uint64_t bytes);
int atx_device_free(void *ptr);
int atx_host_alloc(void **ptr, uint64_t bytes);
int atx_host_free(void *ptr);
int atx_memcpy(void *dst, const void *src, uint64_t bytes, void *stream);
int atx_stream_create(void **stream);
int atx_stream_destroy(void *stream);
int atx_stream_synchronize(void *stream);
int atx_stream_wait_event(void *stream, void *event);
int atx_event_create(void **event);
int atx_event_destroy(void *event);
int atx_event_record(void *event, void *stream);
int atx_event_query(void *event, int *done);
int atx_event_synchronize(void *event);
int atx_event_elapsed(float *ms, void *start, void *end);

#ifdef __cplusplus
}
#endif

#endif /* ATX_H */
