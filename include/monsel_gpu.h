/*
 * monsel_gpu.h -- C-ABI of the MI355X electron-transport engine.
 *
 * This is the drop-in boundary for CudaMONSEL's transport hot path.  The
 * reference runs the path through pure-virtual C++ interfaces that are called
 * per step on the host:
 *   IMaterialScatterModel  gov/nist/microanalysis/NISTMonte/IMaterialScatterModel.cuh:8-18
 *   ScatterMechanism       gov/nist/nanoscalemetrology/JMONSEL/ScatterMechanism.cuh:8-14
 *   SlowingDownAlg         gov/nist/nanoscalemetrology/JMONSEL/SlowingDownAlg.cuh:8-15
 *   BarrierScatterMechanism gov/nist/nanoscalemetrology/JMONSEL/BarrierScatterMechanism.cuh:7-11
 *   Shape / NormalShape    gov/nist/microanalysis/NISTMonte/Shape.cuh:8-15,
 *                          gov/nist/nanoscalemetrology/JMONSEL/NormalShape.cuh:9-19
 *   RegionBase             gov/nist/microanalysis/NISTMonte/RegionBase.cuh:29-31
 *   ElectronGun            gov/nist/microanalysis/NISTMonte/ElectronGun.cuh:8-16
 *   ActionListener/BackscatterStats gov/nist/microanalysis/NISTMonte/BackscatterStats.cu:140-181
 * Here the host serialises that object graph once (monsel_scene_desc, SI
 * units, the reference's own constructor parameters) and the whole trajectory
 * loop -- MonteCarloSS::runMultipleTrajectories
 * (gov/nist/microanalysis/NISTMonte/MonteCarloSS.cu:222-231) -- runs on the GPU.
 *
 * Conventions
 *   - every function returns 0 (MONSEL_OK) or a negative monsel_status;
 *     monsel_last_error() returns a thread-local message for the last failure;
 *   - no C++ exception crosses this boundary;
 *   - the library owns all device memory, the caller owns all host memory;
 *   - descriptors are copied at upload time, so the caller may free them after
 *     monsel_scene_upload / monsel_job_submit return;
 *   - a context is bound to one device; calls on one context are serialised
 *     internally and may come from any host thread.
 */
#ifndef MONSEL_GPU_H
#define MONSEL_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MONSEL_ABI_VERSION 3u

typedef enum monsel_status {
  MONSEL_OK = 0,
  MONSEL_E_INVALID = -1,     /* malformed descriptor or argument */
  MONSEL_E_UNSUPPORTED = -2, /* mechanism/shape combination the kernels do not implement */
  MONSEL_E_NOMEM = -3,       /* device allocation failed */
  MONSEL_E_DEVICE = -4,      /* HIP runtime error (no device, launch failure, ...) */
  MONSEL_E_NOTFOUND = -5,    /* unknown scene or job id */
  MONSEL_E_TIMEOUT = -6,     /* monsel_job_wait timed out */
  MONSEL_E_NOTREADY = -7,    /* job still running */
  MONSEL_E_CANCELLED = -8,   /* job ended by monsel_job_cancel: no results */
  MONSEL_E_TOOLARGE = -9     /* job exceeds the fixed device buffer of monsel_run_static */
} monsel_status;

/* Reference-compatibility switches ("RefCompat").  Each bit reproduces one
 * deterministic defect of the reference; clear it to get the corrected
 * behaviour.  MONSEL_COMPAT_REFERENCE sets all of them.  (The left-corner
 * defect of NShapes.cu:140 is chosen when the line shape is built:
 * NShapes::createLine(..., referenceCompat).) */
enum {
  /* ExpQMBarrierSM.cu:111 overwrites the surface normal with the electron
   * direction: every barrier hit is at normal incidence. */
  MONSEL_COMPAT_BARRIER_NORMAL = 1u << 0,
  /* Algorithm.cu:5-14 + Histogram.cu:73-79: values between bin edges are
   * filed one bin high (binarySearch returns the insertion point). */
  MONSEL_COMPAT_HISTOGRAM_BIN = 1u << 1,
  /* NULagrangeInterpolation.cu:236: d3 builds its outer abscissae from the
   * start of the table instead of index0. */
  MONSEL_COMPAT_D3_ABSCISSAE = 1u << 2,
  MONSEL_COMPAT_REFERENCE = 0x7u,
  MONSEL_COMPAT_FIXED = 0x0u
};

#define MONSEL_MAX_ELEMENTS 8
#define MONSEL_MAX_CORE 8
#define MONSEL_MAX_MECH 6
#define MONSEL_MOTT_NE 61  /* NISTMottRS.cu:23 SPWEM_LEN */
#define MONSEL_MOTT_NQ 201 /* NISTMottRS.cu:24 X1_LEN */

/* ---- materials: SEmaterial (gov/nist/nanoscalemetrology/JMONSEL/SEmaterial.cu) */
typedef struct monsel_material_desc {
  int32_t n_elements; /* 0 = vacuum */
  int32_t Z[MONSEL_MAX_ELEMENTS];
  double weight_fraction[MONSEL_MAX_ELEMENTS]; /* normalised (Composition::weightFraction(el, true)) */
  double atomic_weight[MONSEL_MAX_ELEMENTS];   /* g/mol, Element::getAtomicWeight (FP32 table values) */
  double density;            /* kg/m^3 */
  double workfunction;       /* J  (SEmaterial::getWorkfunction) */
  double energy_cb_bottom;   /* J  (SEmaterial::getEnergyCBbottom) */
  double bandgap;            /* J  (SEmaterial::getBandgap) */
  int32_t n_core;
  int32_t pad0;
  double core_energy[MONSEL_MAX_CORE]; /* J, ascending (SEmaterial::getCoreEnergyArray) */
} monsel_material_desc;

/* ---- scatter mechanisms (ScatterMechanism subclasses) */
typedef enum monsel_mech_kind {
  /* SelectableElasticSM(mat, NISTMottRS::Factory{,100,100Lin}); p[] unused,
   * method = NISTMottRS method 1..3 (NISTMottRS.cu:76-91). */
  MONSEL_MECH_NISTMOTT_ELASTIC = 1,
  /* SelectableElasticSM with BrowningEmpiricalCrossSection for every element. */
  MONSEL_MECH_BROWNING_ELASTIC = 2,
  /* SelectableElasticSM with ScreenedRutherfordScatteringAngle. */
  MONSEL_MECH_RUTHERFORD_ELASTIC = 3,
  /* FittedInelSM(mat, energySEgen, sdAlg): p[0] = energySEgen (J),
   * p[1] = breakE of sdAlg (J), p[2] = sdAlg kind (monsel_csd_kind; ZeroCSD
   * gives a zero rate).  sdAlg acts on the MSM's material. */
  MONSEL_MECH_FITTED_INELASTIC = 4,
  /* GanachaudMokraniPolaronTrapSM(prefactor, extinction):
   * p[0] = prefactor (1/m), p[1] = extinction (1/J). */
  MONSEL_MECH_POLARON_TRAP = 5,
  /* GanachaudMokraniPhononInelasticSM(ratemultiplier, phononE, T, eps0, epsInf):
   * p[0..4] in that order (phononE in J, T in K). */
  MONSEL_MECH_PHONON = 6,
  /* TabulatedInelasticSM(mat, methodSE, tables[4], energyOffset):
   * method = methodSE (1..3), table[0..3] = IIMFP, reduced deltaE, theta, SE E0
   * (indices into monsel_scene_desc.nu_tables, -1 if unused),
   * p[0] = energyOffset (J), p[1] = minEgenSE (J), p[2] = rateMult,
   * p[3] = E0fromDispersion (0/1). */
  MONSEL_MECH_TABULATED_INELASTIC = 7
} monsel_mech_kind;

typedef struct monsel_mech_desc {
  int32_t kind;
  int32_t method;
  int32_t table[4];
  double p[8];
} monsel_mech_desc;

typedef enum monsel_csd_kind {
  MONSEL_CSD_ZERO = 0,            /* ZeroCSD */
  MONSEL_CSD_JOY_LUO_NIEMINEN = 1 /* JoyLuoNieminenCSD(mat, breakE) */
} monsel_csd_kind;

typedef enum monsel_barrier_kind {
  MONSEL_BARRIER_NONE = 0,          /* NullMaterialScatterModel: region hand-off only */
  MONSEL_BARRIER_EXPQM_CLASSICAL = 1, /* ExpQMBarrierSM(mat) */
  MONSEL_BARRIER_EXPQM = 2            /* ExpQMBarrierSM(mat, lambda) */
} monsel_barrier_kind;

/* MONSEL_MaterialScatterModel (NM/JMONSEL/MONSEL_MaterialScatterModel.cu) or,
 * with is_null = 1, NullMaterialScatterModel (MA/NISTMonte/NullMaterialScatterModel.cu). */
typedef struct monsel_msm_desc {
  int32_t material;     /* index into materials */
  int32_t is_null;
  int32_t n_mech;       /* mechanisms in addScatterMechanism order */
  int32_t csd_kind;
  monsel_mech_desc mech[MONSEL_MAX_MECH];
  double csd_break_e;   /* J, JoyLuoNieminenCSD breakE */
  int32_t barrier_kind;
  int32_t pad0;
  double barrier_lambda; /* m, ExpQMBarrierSM lambda */
  double min_e_for_tracking; /* J; -INFINITY allowed (vacuum) */
} monsel_msm_desc;

/* ---- shapes */
typedef enum monsel_shape_kind {
  MONSEL_SHAPE_PLANES = 1,       /* NormalMultiPlaneShape: planes[first_plane .. +n_planes) */
  MONSEL_SHAPE_CYLINDER = 2,     /* NormalCylindricalShape(end0, end1, radius) */
  MONSEL_SHAPE_INTERSECTION = 3, /* NormalIntersectionShape(a, b) */
  MONSEL_SHAPE_UNION = 4,        /* NormalUnionShape(a, b) */
  MONSEL_SHAPE_SPHERE = 5        /* Sphere(center, radius): chamber only */
} monsel_shape_kind;

typedef struct monsel_plane_desc {
  double normal[3]; /* outward; inside is (p - point).normal <= 0 (MultiPlaneShape.cu:23-27) */
  double point[3];
} monsel_plane_desc;

typedef struct monsel_shape_desc {
  int32_t kind;
  int32_t a, b;              /* child shapes of INTERSECTION / UNION */
  int32_t first_plane, n_planes;
  int32_t pad0;
  double end0[3], end1[3];   /* CYLINDER (m) */
  double radius;             /* CYLINDER / SPHERE (m) */
  double center[3];          /* SPHERE (m) */
} monsel_shape_desc;

/* Region(parent, msm, shape) (MA/NISTMonte/RegionBase.cu:193-200).  Region 0
 * is the chamber (parent = -1).  Children are visited in ascending index. */
typedef struct monsel_region_desc {
  int32_t parent;
  int32_t shape;
  int32_t msm;
  int32_t pad0;
} monsel_region_desc;

/* NIST SRD-64 table of one element (NISTMottScatteringAngle::loadData). */
typedef struct monsel_mott_table {
  int32_t Z;
  int32_t pad0;
  const double* spwem; /* [61] total cross-section in Bohr radius^2 */
  const double* x1;    /* [61][201] angular quantiles */
} monsel_mott_table;

/* NUTableInterpolation file contents (NM/JMONSEL/NUTableInterpolation.cu:117-228). */
typedef struct monsel_nu_table {
  int32_t dim;          /* 1..4 */
  int32_t n[4];
  int32_t pad0;
  const double* x[4];   /* monotonic abscissae per dimension */
  const double* values; /* prod(n) values, last index fastest */
} monsel_nu_table;

typedef struct monsel_scene_desc {
  uint32_t abi_version; /* MONSEL_ABI_VERSION */
  uint32_t compat_flags;
  int32_t n_materials;
  int32_t n_msm;
  int32_t n_shapes;
  int32_t n_planes;
  int32_t n_regions;
  int32_t n_mott;
  int32_t n_nu_tables;
  int32_t pad0;
  double chamber_radius; /* m, MonteCarloSS::ChamberRadius = 0.1 */
  const monsel_material_desc* materials;
  const monsel_msm_desc* msms;
  const monsel_shape_desc* shapes;
  const monsel_plane_desc* planes;
  const monsel_region_desc* regions;
  const monsel_mott_table* mott;
  const monsel_nu_table* nu_tables;
} monsel_scene_desc;

/* ---- jobs: GaussianBeam + runMultipleTrajectories per pixel */
typedef struct monsel_job_desc {
  double beam_energy;     /* J (GaussianBeam::setBeamEnergy) */
  double beam_width;      /* m, Gaussian sigma (GaussianBeam.cu:59) */
  int32_t n_pixels;
  int32_t primaries_per_pixel;
  const double* pixel_center; /* [n_pixels][3] beam centres (m), GaussianBeam::setCenter */
  uint64_t seed;
  int32_t n_energy_bins;  /* BackscatterStats(mcss, nEnergyBins): bins over [0, E0 in eV] */
  int32_t max_steps;      /* per-electron watchdog; <= 0 selects 1000000 */
  double se_cutoff_ev;    /* SE yield = back-escapes in bins j < (int)(cutoff / binwidth) */
  uint64_t primary_offset; /* first primary index (for splitting one job) */
  int64_t log_detected;    /* BackscatterStats::setLogDetected (BackscatterStats.cu:166-168): keep up to this
                            * many detected-electron records (monsel_job_read_detected); 0 = none */
} monsel_job_desc;

/* Caller-owned result buffers; any pointer may be NULL.  Histogram slot
 * layout follows Histogram.cu:25 (slot 0 underrange, 1..nbins, nbins+1 overrange). */
typedef struct monsel_tallies {
  uint64_t* back;       /* [n_pixels] escapes with elevation >= pi/2 */
  uint64_t* fwd;        /* [n_pixels] escapes with elevation < pi/2 */
  uint64_t* se;         /* [n_pixels] back-escapes in the SE bins */
  double* se_sq;        /* [n_pixels] sum over primaries of (SE count)^2 */
  double* bse_sq;       /* [n_pixels] sum over primaries of (back - SE count)^2 */
  double* back_energy;  /* [n_pixels] sum of exit energies of back-escapes (eV) */
  uint64_t* back_energy_hist; /* [n_energy_bins + 2] */
  uint64_t* fwd_energy_hist;  /* [n_energy_bins + 2] */
  uint64_t* elevation_hist;   /* [180 + 2] over [0, pi] */
  uint64_t* azimuth_hist;     /* [360 + 2] over [0, 2 pi] */
} monsel_tallies;

/* One detected electron (BackscatterStats::Datum, BackscatterStats.cu:40-48,
 * logged at :166-168): which electron, its step count, exit energy, the
 * point where it reached the chamber wall and its direction there. */
typedef struct monsel_detected {
  uint64_t electron;     /* lineage of the electron: unique within its primary's cascade */
  uint32_t primary;      /* primary index (the cascade it belongs to) */
  uint32_t step;         /* Electron::getStepCount at detection */
  double energy_ev;      /* exit energy (eV) */
  double position[3];    /* the tally point on the chamber sphere (m) */
  double theta, phi;     /* direction: polar angle from +z, azimuth (Electron::getTheta/getPhi) */
} monsel_detected;

typedef struct monsel_job_stats {
  uint64_t primaries;
  uint64_t electrons;     /* primaries + tracked secondaries that took a step */
  uint64_t steps;         /* takeStep calls */
  uint64_t escapes;
  uint64_t watchdog_drops;
  uint64_t queue_overflow; /* secondaries lost to a full generation buffer */
  int32_t generations;
  int32_t transport_launches; /* launches of the transport kernel */
  double device_ms;       /* transport + reduction kernels */
  double total_ms;        /* submit .. results on device */
  double transport_ms;    /* sum of transport-kernel durations (HIP events) */
} monsel_job_stats;

/* Launch shape; zero fields select defaults. */
typedef struct monsel_tuning {
  int32_t block_size;        /* threads per workgroup (multiple of 64, <= 256) */
  int32_t blocks_per_cu;     /* persistent workgroups per CU */
  int32_t max_generations;   /* transport launches per job before its electrons are dropped */
  int32_t queue_capacity;    /* electrons per generation buffer (0 = from the device memory) */
  int32_t max_jobs_in_flight;   /* jobs sharing the transport launches (0 = 1024) */
  int32_t primaries_per_launch; /* new primaries admitted per launch (0 = 262,144, or the queue capacity
                                 * if smaller) */
  int32_t steps_per_launch;     /* regroup interval: steps an electron takes per launch (0 = unlimited) */
  int32_t admission_order;      /* new primaries go first to: 0 the oldest job, 1 the job with fewest left */
  int32_t sort_queue;           /* 0: sort each launch's queue by region and energy class; -1: keep spawn order */
  int32_t geometry_batch;       /* lanes of a wave whose step may reach a boundary that wait before the wave
                                 * runs findEndOfStep for them (<= 64; 0 = by scene: 8 where a region's own
                                 * and child shapes hold 3 or more primitives, else 1 = no waiting) */
  int32_t spawn_handoff;        /* 0: a new secondary goes straight to an idle lane of the wave that
                                 * spawned it when there is one; -1: every one waits in the queue
                                 * for the next launch */
  int32_t tail_requeue;         /* once a launch's own items are all claimed, a wave down to this many
                                 * live lanes queues them for the next launch and leaves
                                 * (0 = default 32, -1 = never; launches of < 65536 items never do) */
} monsel_tuning;

typedef struct monsel_ctx_s* monsel_ctx;
typedef uint64_t monsel_scene_id;
typedef uint64_t monsel_job_id;

const char* monsel_last_error(void);
int monsel_abi_version(void);

/* Context on one HIP device ordinal. tuning may be NULL. */
int monsel_create(int device_ordinal, const monsel_tuning* tuning, monsel_ctx* out);
int monsel_destroy(monsel_ctx ctx);
int monsel_device_info(monsel_ctx ctx, int* n_cu, size_t* hbm_bytes, char* name, size_t name_len);
/* Cumulative counters of a context: transport launches, electron steps of
 * completed jobs and transport-kernel time (HIP events on the context's
 * stream).  No reference counterpart: for throughput measurement when jobs
 * share launches (a job's own transport_ms counts every launch it was in). */
int monsel_ctx_counters(monsel_ctx ctx, uint64_t* launches, uint64_t* steps, double* transport_ms);

/* Scenes: an immutable, device-resident copy of the object graph and tables. */
int monsel_scene_footprint(const monsel_scene_desc* desc, size_t* hbm_bytes);
int monsel_scene_upload(monsel_ctx ctx, const monsel_scene_desc* desc, monsel_scene_id* out);
/* Drops the context's reference to a scene; jobs already submitted on it
 * keep it (it is freed when the last of them ends). */
int monsel_scene_release(monsel_ctx ctx, monsel_scene_id scene);

/* Jobs.  submit enqueues the work on the context's stream and returns;
 * wait/poll complete it; read copies tallies into caller buffers. */
int monsel_job_footprint(monsel_ctx ctx, const monsel_job_desc* job, size_t* hbm_bytes);
int monsel_job_submit(monsel_ctx ctx, monsel_scene_id scene, const monsel_job_desc* job, monsel_job_id* out);
int monsel_job_poll(monsel_ctx ctx, monsel_job_id job, int* done);
int monsel_job_wait(monsel_ctx ctx, monsel_job_id job, int64_t timeout_us);
int monsel_job_read(monsel_ctx ctx, monsel_job_id job, monsel_tallies* out);
int monsel_job_stats_get(monsel_ctx ctx, monsel_job_id job, monsel_job_stats* out);
int monsel_job_release(monsel_ctx ctx, monsel_job_id job);
/* Detected-electron records of a completed job that asked for them
 * (monsel_job_desc.log_detected > 0), in no particular order: copies up to
 * max of them into out and sets *n_detected to the number the job detected
 * (more than it kept when its log_detected was too small). */
int monsel_job_read_detected(monsel_ctx ctx, monsel_job_id job, monsel_detected* out, int64_t max,
                             int64_t* n_detected);
/* Ends a job early.  No reference counterpart (runMultipleTrajectories,
 * MonteCarloSS.cu:222-231, always runs to its end); a serving loop uses it to
 * drop a request that timed out.  A job still waiting for admission ends at once; a running one
 * launches no further primaries, its queued electrons are dropped at the next
 * transport launch, and then its slot, device buffers and queue reservation
 * are freed.  Other jobs sharing those launches are unaffected (their results
 * do not depend on the mix).  wait/read on a cancelled job return
 * MONSEL_E_CANCELLED; release it as any other.  Cancelling a job that has
 * already completed changes nothing (returns MONSEL_OK). */
int monsel_job_cancel(monsel_ctx ctx, monsel_job_id job);

/* Convenience: submit + wait + read + release. */
int monsel_run(monsel_ctx ctx, monsel_scene_id scene, const monsel_job_desc* job,
               monsel_tallies* out, monsel_job_stats* stats);

/* monsel_run without device allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (tight serving loops,
 * code that forbids hipMalloc).  The same job on the same kernels, with the
 * same results as monsel_run, run on the calling thread; every device buffer
 * beyond the caller's inputs (job, pixel centres, the uploaded scene) and
 * outputs (tallies, stats) lives in one static device array of 128 MiB per
 * GPU, held from the library's load on that GPU until the process ends.
 * The tallies of a job take 16 B per primary + 96 B per pixel + 8 B per
 * histogram bin; the rest holds the generation queues (148 B per queued
 * electron, at least max(4096, 16 + E0 x secondaries per eV)); a larger job
 * returns MONSEL_E_TOOLARGE.  Calls on one GPU share the array (which keeps
 * the previous call's contents; the tallies are cleared by each call) and
 * would race, so the library runs them one at a time per GPU.  Its kernels
 * share the context's stream with jobs submitted to the context.  Calls no
 * hipMalloc, hipFree or other allocating routine. */
int monsel_run_static(monsel_ctx ctx, monsel_scene_id scene, const monsel_job_desc* job,
                      monsel_tallies* out, monsel_job_stats* stats);

/* ---- probes: one device evaluation per input, for parity tests.  Inputs and
 * outputs are host arrays; energies in eV, lengths in nm. */

/* MONSEL_MaterialScatterModel::setCache (MONSEL_MaterialScatterModel.cu:44-62):
 * out[i*(n_mech+1) + k] = scatterRate of mechanism k at energies[i] (1/nm),
 * out[... + n_mech] = total. */
int monsel_probe_rates(monsel_ctx ctx, monsel_scene_id scene, int msm, int n,
                       const float* energies_ev, float* out);

/* RandomizedScatter::randomScatteringAngle of element slot `element` of the
 * MSM's elastic mechanism (NISTMottRS.cu:121-154, BrowningEmpiricalCrossSection.cu:31-45)
 * with injected uniforms u[2*i], u[2*i+1]; out = scattering angle (rad). */
int monsel_probe_elastic_angle(monsel_ctx ctx, monsel_scene_id scene, int msm, int element, int n,
                               const float* energies_ev, const float* u, float* out);

/* SlowingDownAlg::compute (JoyLuoNieminenCSD.cu:91-121): energy change (eV)
 * over path length len_nm[i] at energies_ev[i]. */
int monsel_probe_csd(monsel_ctx ctx, monsel_scene_id scene, int msm, int n,
                     const float* len_nm, const float* energies_ev, float* out);

/* RegionBase::findEndOfStep (RegionBase.cu:74-119) for a segment p0 -> p1
 * (nm) starting in `region`: out_u[i] = fraction of the segment to the first
 * boundary (> 1 if none), out_next[i] = next region (-1 = left the chamber),
 * out_normal[3*i] = outward normal of the crossed surface (0 if none). */
int monsel_probe_find_end_of_step(monsel_ctx ctx, monsel_scene_id scene, int region, int n,
                                  const float* p0, const float* p1, float* out_u, int* out_next,
                                  float* out_normal);

/* Lower bound (nm) on the distance from pos to every surface findEndOfStep
 * tests from `region` (its shape and its children's shapes).  No reference
 * counterpart: MonteCarloSS::takeStep (MonteCarloSS.cu:107-163) runs
 * findEndOfStep on every step, and the transport kernel skips it for a step
 * shorter than this bound. */
int monsel_probe_safety(monsel_ctx ctx, monsel_scene_id scene, int region, int n, const float* pos, float* out);

/* RegionBase::containingSubRegion from the chamber (RegionBase.cu:50-60). */
int monsel_probe_locate(monsel_ctx ctx, monsel_scene_id scene, int n, const float* pos, int* out_region);

/* NUTableInterpolation::interpolate (NUTableInterpolation.cu:98-115) of
 * scene table `table` at x[i*dim ..] with the given order. */
int monsel_probe_nu_table(monsel_ctx ctx, monsel_scene_id scene, int table, int order, int n,
                          const float* x, float* out);

/* ScatterMechanism::scatter of mechanism slot `mech` of MSM `msm`
 * (SelectableElasticSM.cu:49-67, FittedInelSM.cu:13-18,
 * GanachaudMokraniPolaronTrapSM.cu:15-20, GanachaudMokraniPhononInelasticSM.cu:19-44,
 * TabulatedInelasticSM.cu:209-434) for an electron at energies_ev[i] after a
 * step from prev_ev[i] (the elastic choice uses the pre-step energy), unit
 * direction dir[3*i], with the uniforms u[6*i .. 6*i+5] taken in the order the
 * reference's scatter() draws them.  out[11*i ..]: direction (3), energy (eV),
 * secondary flag, secondary energy (eV), secondary direction (3), trapped flag,
 * 0. */
int monsel_probe_scatter(monsel_ctx ctx, monsel_scene_id scene, int msm, int mech, int n, const float* prev_ev,
                         const float* energies_ev, const float* dir, const float* u, float* out);

/* ExpQMBarrierSM::barrierScatter (ExpQMBarrierSM.cu:98-219) from region
 * `region_from` into `region_to` for unit direction dir[3*i], outward barrier
 * normal normal[3*i] (as resolved from the struck shape by
 * ExpQMBarrierSM.cu:77-96), energy energies_ev[i] and transmission uniform
 * u[i].  out[6*i ..]: direction (3), energy (eV), region after, nudge sign
 * (+1 along the normal, -1 against it). */
int monsel_probe_barrier(monsel_ctx ctx, monsel_scene_id scene, int region_from, int region_to, int n,
                         const float* dir, const float* normal, const float* energies_ev, const float* u, float* out);

/* The device's Philox4x32-10 (the generator that replaces Math2::random,
 * gov/nist/microanalysis/Utility/Math2.cu:1328-1336, with per-electron
 * streams): out[4i..4i+3] = Philox4x32-10(counter ctr[4i..4i+3], key
 * key[2i..2i+1]); out_u (may be NULL) receives the FP32 uniforms in (0, 1)
 * the transport kernel makes of those words. */
int monsel_probe_philox(monsel_ctx ctx, int n, const uint32_t* ctr, const uint32_t* key, uint32_t* out,
                        float* out_u);

/* Histogram::bin (gov/nist/microanalysis/Utility/Histogram.cu:73-79 with
 * Amphibian/Algorithm.cu:5-14) as the escape tallies compute it: bin index
 * of val[i] over nbins bins of [min, max], -1 = underrange, nbins =
 * overrange; compat != 0 selects the reference's insertion-point result. */
int monsel_probe_hist_bin(monsel_ctx ctx, int n, const double* val, double min, double max, int nbins, int compat,
                          int* out);

/* GaussianBeam::createElectron (gov/nist/microanalysis/NISTMonte/GaussianBeam.cu:52-65)
 * for primaries 0 .. n-1 of `job` as the transport kernel starts them (same
 * random stream, same pixel order): out[4i..4i+3] = position (nm) and the
 * region containing it (-1 = outside the chamber). */
int monsel_probe_primaries(monsel_ctx ctx, monsel_scene_id scene, const monsel_job_desc* job, int n, float* out);

#ifdef __cplusplus
}
#endif
#endif /* MONSEL_GPU_H */
