/*
 * kerr_trace.h -- C ABI of the MI355X Kerr photon tracer (libkerr_trace.so).
 *
 * A drop-in for the Black_simulation per-pixel photon trace: camera pose, spin,
 * disk and escape radii and image size go in; per-ray termination class,
 * asymptotic sky angles, disk-hit radius and redshift factor g come out.
 *
 * The reference has no plugin or FFI layer; the trace is reached through three
 * in-process C++ calls, which these entry points replace:
 *   kt_trace_views  <- eht::render            (src/eht.h:140, src/eht.cpp:321-347)
 *                      and the trace pass of Renderer::render (src/renderer.cpp:283-379,
 *                      shaders/trace.frag:1490-1800)
 *   kt_trace_rays   <- eht::traceRay          (src/eht.h:130, src/eht.cpp:260-319)
 *                      trajectory::traceRay   (src/trajectory.h:118-119) for a look list
 *   kt_horizon_radius, kt_isco_radius, kt_photon_orbit_radius, kt_circular_omega
 *                   <- kerr::horizonRadius, iscoRadius, photonOrbitRadius,
 *                      circularOmega          (src/kerr.h:39-45, 71)
 *   kt_critical_curve, kt_critical_radius
 *                   <- eht::criticalCurve, eht::criticalRadius (src/eht.h:66-68)
 *   kt_eht_render, kt_eht_trace_rays, kt_eht_flux, kt_eht_calibrate
 *                   <- eht::render, eht::traceRay (with the light), eht::flux,
 *                      eht::calibrate         (src/eht.h:130-152)
 *   KT_CAMERA_OBSERVER views, kt_geodesic_four_velocity
 *                   <- Camera::frame          (src/camera.cpp:23-40), kerr::fourVelocity
 *   kt_particles_advance, kt_particle_from_zamo
 *                   <- the debris update      (shaders/particles_update.comp:207-255),
 *                      kerr::stepCoordinateTime, kerr::launchFromZamo (src/kerr.cpp)
 *
 * Conventions
 *   - Units G = c = M = 1.  Angles in radians except fov_deg.
 *   - Sky angles are Boyer-Lindquist (theta, phi) of the ray's asymptotic direction
 *     at r -> infinity, traced backward from the camera (phi in [0, 2 pi)).
 *   - Outputs are structure-of-arrays, views concatenated in order, each view
 *     row-major from the top row (beta = +half_height), as eht::Image.
 *   - Every call returns a kt_status; per-ray anomalies are classes, never errors.
 *     No C++ exception crosses this boundary.
 *   - A context is externally synchronised: one call at a time per context, on the
 *     caller's HIP stream (NULL = the default stream).  Several contexts may coexist.
 */
#ifndef KERR_TRACE_H
#define KERR_TRACE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum kt_status {
    KT_OK = 0,
    KT_EINVAL = 1,        /* bad argument (message in kt_last_error_string) */
    KT_EHIP = 2,          /* a HIP runtime call failed                      */
    KT_ECANCELLED = 3,    /* the cancel flag was set                        */
    KT_ENOMEM = 4,        /* device allocation failed                       */
    KT_EUNSUPPORTED = 5,  /* feature not available (e.g. no surrogate)      */
    KT_ENODEVICE = 6      /* no usable GPU                                  */
} kt_status;

/* per-ray class codes (low 3 bits of cls) -- trajectory::RayEnd (src/trajectory.h:85) */
enum {
    KT_ESCAPED = 0,    /* RayEnd::Sky                                   */
    KT_CAPTURED = 1,   /* RayEnd::Horizon: r < capture_factor * r_+       */
    KT_DISK = 2,       /* RayEnd::Disk                                  */
    KT_UNFINISHED = 3  /* RayEnd::Unfinished: step budget exhausted      */
};
/* flag bits OR-ed into cls */
enum {
    KT_FLAG_CLOSED_FORM = 0x08,  /* answered by the closed-form solution (0 steps) */
    KT_FLAG_FP64_RESTEP = 0x10,  /* mixed path: at least one FP64 re-step          */
    KT_FLAG_SURROGATE = 0x20,    /* answered by the learned surrogate              */
    KT_FLAG_NONFINITE = 0x40     /* the state went non-finite: counted as captured */
};

enum { KT_CAMERA_BARDEEN = 0, KT_CAMERA_PINHOLE = 1, KT_CAMERA_OBSERVER = 3 };
enum { KT_PREC_FP64 = 0, KT_PREC_MIXED = 1, KT_PREC_FP32 = 2 };
enum { KT_MODE_PERSISTENT = 0, KT_MODE_ONESHOT = 1, KT_MODE_AUTO = 2 /* the tuned one, else persistent */ };

/* One camera view.  Bardeen views (eht::View, src/eht.h:114-121) trace the image
 * plane of a camera at (r_cam, theta_cam = inclination, phi_cam): pixel centres
 * alpha = -half_width + (col + 0.5) 2 half_width / width, beta likewise from the top
 * (src/eht.cpp:332-334).  Pinhole views are the simulator's camera
 * (src/camera.h:20-29: pose, yaw/pitch/roll, vertical fov) in the ZAMO frame.
 * Observer views are the same camera carried by any observer (Camera::frame,
 * src/camera.cpp:23-40, and kerr::observerFrame, src/kerr.cpp:432-456): the 4-velocity
 * u_obs (ingoing Kerr-Schild components, normalised here), or, when u_obs is all zero,
 * the ZAMO boosted by `drift` (its 3-velocity along the ZAMO's e_r, e_theta, e_phi,
 * |drift| < 1).  With u_obs the camera may sit inside the horizon (above r-): rays that
 * reach the horizon from the far side are captured at once (src/trajectory.cpp:234-241),
 * the others are followed out through it.  For observer views phi_cam is the Kerr-Schild
 * phi, as in the simulator; they trace in FP64 only and cannot share a call with other
 * kinds.  Disk fields follow trajectory::RayScene (src/trajectory.h:74-83).  Every real
 * field must be finite: a NaN or infinity anywhere is KT_EINVAL.  spin lies in [-1, 1.5]:
 * past 1 a naked singularity (src/kerr.cpp:117), traced for Bardeen and pinhole views by the
 * integrator; it has no horizon, so capture_factor x r+ = 0 and a ray is captured where it
 * passes through the ring's disk to r < 0 (as eht::traceRay does, src/eht.cpp:285). */
typedef struct kt_view {
    int camera_kind;
    int width, height;
    double spin;
    double r_cam, theta_cam, phi_cam;
    double half_width, half_height;
    double yaw, pitch, roll, fov_deg;
    double disk_inner, disk_outer;  /* disk_inner <= 0: prograde/retrograde ISCO */
    int disk_prograde, disk_enabled;
    double escape_radius;           /* escape when r > max(escape_radius, 1.5 r_cam), outgoing */
    double capture_factor;          /* captured when r < capture_factor * r_+ (e.g. 1.01) */
    double u_obs[4];                /* KT_CAMERA_OBSERVER: u^t, u^r, u^theta, u^phi (KS)  */
    double drift[3];                /* KT_CAMERA_OBSERVER with u_obs = 0: ZAMO boost      */
    /* KT_CAMERA_OBSERVER: the observer's frame itself (as uCamPhoton carries Camera::frame,
     * src/renderer.cpp:345-357), e_a^mu at 4 a + mu: row 0 the 4-velocity, rows 1-3 the r,
     * theta and phi axes of kerr::Frame (kerr.h), mu = (t, r, theta, phi) ingoing Kerr-Schild.
     * It must be orthonormal (g(e_a, e_b) = diag(-1, 1, 1, 1) to 1e-8) and future-directed.
     * All zero: the frame is built from u_obs or drift (kt_observer_frame). */
    double frame[16];
} kt_view;

typedef struct kt_trace_opts {
    int precision;          /* KT_PREC_*                                            */
    int mode;               /* KT_MODE_* (default KT_MODE_AUTO)                      */
    int use_surrogate;      /* 1: surrogate answers confident rays, others re-queued */
    int max_steps;          /* per ray, accepted + rejected (default 200000)        */
    double rtol, atol;      /* FP64 DP5 tolerances (default 1e-10, 1e-12)           */
    double rtol32, atol32;  /* FP32 bulk tolerances (default 1e-6, 1e-7)            */
    double restep_drift;    /* mixed: FP64 re-step when |H| drift > this (1e-6)     */
    int block;              /* threads per block (0 = tuned)                        */
    int waves_per_cu;       /* persistent grid size (0 = tuned)                     */
    int steps_per_dequeue;  /* persistent refill period (0 = tuned)                 */
    double gate_prob;       /* surrogate: min class probability (0.999)             */
    double gate_ring;       /* surrogate: min |rho/rho_ring - 1| (0.01)             */
    int waves_per_simd;     /* register budget of the persistent kernel: 1-4 (0 = tuned) */
    /* kt_eht_render, kt_eht_trace_rays and kt_eht_calibrate (as eht::render's and
     * eht::calibrate's cancel, src/eht.h:140-153): a host flag read between slices of the
     * image (and between calibration renders).  Once it reads nonzero no further slice is
     * traced, the rays not traced are left zero (as eht::render leaves its rows) and the
     * call returns KT_ECANCELLED; calibrate still writes its current estimate.  A cancellable
     * render runs slice by slice, synchronising the stream after each.  NULL: not
     * cancellable.  The other calls ignore it. */
    const volatile int* cancel;
    /* 1: the closed-form solution of Carter's equations (Legendre elliptic integrals of the
     * Mino-time quadratures, FP64) answers the rays of Bardeen planes, pinhole cameras (the
     * static ZAMO's, or any observer's outside the horizon) and explicit rays that it covers
     * -- class, disk radius and g, escaping rays' sky angles -- with 0 steps and
     * KT_FLAG_CLOSED_FORM; the rays it declines (an escape through a polar vortical orbit or
     * through the pole, a turning point at the capture radius, a crossing within 1e-9 of a
     * disk edge, outgoing rays whose radial potential has no real root or a complex pair
     * too near the path for the escape's quadrature, launches with R or Theta negative at
     * the camera, observer cameras inside the horizon, spins
     * past 1) go to the integrator at `precision`.  With use_surrogate, the network's
     * gate runs first and the closed form replaces the values of the rays it answered as
     * escaped or disk hits (re-queueing those it declines or classes otherwise).  Not with
     * the ray report.  0: off (the default).
     * 2: the same with the elliptic integrals and Jacobi functions in FP32 (the roots of the
     * radial potential and the class decisions they make stay FP64): disk radius and g to
     * ~1e-6; a ray whose crossing lies within 1e-5 of a disk edge, or that escapes, is
     * solved again in FP64. */
    int closed_form;
} kt_trace_opts;

/* SoA outputs; any pointer may be NULL.  on_device: 1 = device pointers
 * (written on the call's stream), 0 = host pointers (the call synchronises). */
typedef struct kt_rays_out {
    int8_t* cls;
    double* theta_inf;
    double* phi_inf;
    double* r_hit;
    double* g;
    int32_t* steps;
    int on_device;
    /* optional FP32 outputs, written beside or instead of the doubles above (a field may
     * be given in either width or both): each the FP64 result rounded once, NaN where it
     * does not apply.  Same placement as the others (on_device). */
    float* theta_inf_f32;
    float* phi_inf_f32;
    float* r_hit_f32;
    float* g_f32;
    /* optional single-ray report, trajectory::RayReport (src/trajectory.h:97-113) as
     * App::inspectPixel prints it (src/app.cpp:776-777); any of them given selects an FP64
     * one-shot trace (KT_PREC_FP64 only, no surrogate) whose other outputs are the same as
     * without them.  Each ray is followed to its event: the disk crossing, r = r_cap or
     * r = r_esc (located on the step's dense output), or its step budget.
     *   turns  angle swept by the direction of the position in Kerr-Schild Cartesian form
     *          (kerr::toCartesian), over 2 pi, up to the event
     *   r_min  least r on the way (camera, event, and radial turning points as roots of R)
     *   delay  ingoing Kerr-Schild coordinate time from the event to the camera
     *   b, q   L/E and Q/E^2 (0 when |E| <= 1e-12, as the reference) */
    double* turns;
    double* r_min;
    double* delay;
    double* b;
    double* q;
} kt_rays_out;

typedef struct kt_device_opts {
    int device;  /* HIP device ordinal */
} kt_device_opts;

typedef struct kt_stats {
    double kernel_ms;          /* device time of the last trace call's kernels (HIP events) */
    double integrator_ms;      /* ... of its integrator kernel alone                        */
    double surrogate_ms;       /* ... of its surrogate inference or closed-form kernels     */
    long long rays;            /* rays traced                                               */
    long long steps;           /* DP5 attempts, summed over rays                            */
    long long fallback_rays;   /* surrogate / closed form: rays re-queued to the integrator */
    long long restep_rays;     /* mixed path: rays with at least one FP64 step              */
    long long fp64_steps;      /* mixed path: DP5 attempts taken in FP64                     */
    int block, waves_per_cu, steps_per_dequeue, grid, waves_per_simd, mode;
} kt_stats;

typedef struct kt_ctx kt_ctx;

void kt_trace_opts_default(kt_trace_opts* opts);
void kt_view_default(kt_view* view);

int kt_create(const kt_device_opts* opts, kt_ctx** ctx);
int kt_destroy(kt_ctx* ctx);
const char* kt_last_error_string(const kt_ctx* ctx);
int kt_last_stats(const kt_ctx* ctx, kt_stats* stats);

/* Rays of all views in one launch (one device work queue). */
int kt_trace_views(kt_ctx* ctx, const kt_view* views, int n_views, const kt_trace_opts* opts,
                   const kt_rays_out* out, void* hip_stream);

/* kt_trace_views for callers that cannot allocate: code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc.  No counterpart in the reference.  It
 * calls no allocating routine: the view table and the work-queue counters live in a fixed
 * device array of each GPU (about 26 KB: KT_NOALLOC_MAX_VIEWS views), whose address
 * kt_create resolves and which is held until the process ends.  Limits, each refused with
 * KT_EINVAL: at most KT_NOALLOC_MAX_VIEWS views, outputs on the device (on_device = 1), no
 * surrogate (its fallback queue is allocated).  The array is shared by every call on one
 * GPU, from any context: calls that overlap on one GPU race and must be serialised by the
 * caller.  Results equal kt_trace_views' for the same arguments; kt_last_stats reports it. */
#define KT_NOALLOC_MAX_VIEWS 64
int kt_trace_views_noalloc(kt_ctx* ctx, const kt_view* views, int n_views, const kt_trace_opts* opts,
                           const kt_rays_out* out, void* hip_stream);

/* Explicit rays of one view: for a Bardeen view `params` holds n (alpha, beta) pairs;
 * for a pinhole view n unit look vectors (x, y, z) along the frame's (r, theta, phi)
 * axes (the `look` of kerr::photonFromFrame, src/kerr.cpp:458).  params is a host
 * pointer unless out->on_device, then a device pointer. */
int kt_trace_rays(kt_ctx* ctx, const kt_view* view, const double* params, long long n,
                  const kt_trace_opts* opts, const kt_rays_out* out, void* hip_stream);

/* The launch-shape autotuner (no reference counterpart: the reference rasterises one
 * fragment per pixel).  Times the trace of `views` under candidate register budgets
 * (waves_per_simd), grid sizes (waves_per_cu), refill periods (steps_per_dequeue) and
 * block sizes by coordinate descent with HIP events, writes the fastest into *opts and
 * remembers it in ctx for later calls whose fields are 0.  Synchronous; the device
 * buffers it uses are the context's own. */
int kt_autotune(kt_ctx* ctx, const kt_view* views, int n_views, kt_trace_opts* opts);

/* A prepared batch: the view table uploaded once and the launch sequence captured in a
 * hipGraph, replayed by kt_plan_launch (no reference counterpart; the batched-views
 * path of a frame loop).  `out` must hold device pointers; they are baked into the
 * graph.  With opts->use_surrogate the graph holds the network, the gate and the
 * integrator over the rays the gate rejected.  A plan is externally synchronised; it owns
 * its work-queue counters and fallback queue, so different plans (with different
 * outputs) may replay concurrently on different streams.  The surrogate's parameters
 * are read at replay time. */
typedef struct kt_plan kt_plan;
int kt_plan_create(kt_ctx* ctx, const kt_view* views, int n_views, const kt_trace_opts* opts,
                   const kt_rays_out* out, kt_plan** plan);
int kt_plan_launch(kt_plan* plan, void* hip_stream);
/* Counters and time of the plan's last replay (waits for it): kernel_ms spans the whole
 * graph; integrator_ms equals it for integrator-only plans and is 0 for surrogate plans. */
int kt_plan_stats(kt_plan* plan, kt_stats* stats);
int kt_plan_destroy(kt_plan* plan);

/* ---- the learned surrogate (no reference counterpart) ----------------------------------
 * An MLP on the matrix cores with FP32 accumulation and FP32 master weights (BF16 operands
 * in training; FP16 after the input layer in the fast network's inference), one of two
 * networks: KT_NET_WIDE, 5-256-256-256-256-8 with SiLU, and KT_NET_FAST,
 * 5-128-128-128-128-8 with hard SiLU (x clamp(x/6 + 1/2, 0, 1)), a quarter of the matrix
 * work and no transcendentals.  It answers a ray when its largest class probability
 * exceeds opts.gate_prob and the ray lies outside opts.gate_ring of the critical curve;
 * every other ray goes to the integrator (kt_trace_views with use_surrogate = 1).
 * Parameters are FP32, layer by layer: W_l (n_in x n_out, row-major) then b_l; their
 * count names the network (kt_surrogate_net_param_count). */
enum { KT_NET_WIDE = 0, KT_NET_FAST = 1 };
typedef struct kt_surrogate_opts {
    long long n_train, n_holdout;  /* 2^24, 2^20                                        */
    int epochs, batch;             /* 192, 65536                                        */
    double lr;                     /* 1e-3, cosine to 0                                 */
    double w_sky, w_disk;          /* loss weights of the sky and disk regressions (30, 1) */
    double band, band_factor;      /* rays within `band` of the critical curve drawn    */
                                   /* `band_factor` x more often (0.02, 4)              */
    unsigned long long seed;
    int label_precision;           /* KT_PREC_* labelling the training set (mixed)      */
    double disk_outer, escape_radius, capture_factor;  /* the scene it learns (20, 1000, 1.01) */
    int verbose;                   /* 1: one progress line per epoch on stderr          */
    int net;                       /* KT_NET_* to train (KT_NET_WIDE)                   */
    double edge_band;              /* training rays hitting the disk within edge_band of */
    int edge_copies;               /* its inner edge get edge_copies jittered neighbours */
                                   /* (0.06, 4; edge_copies < 0: none)                  */
} kt_surrogate_opts;

typedef struct kt_surrogate_report {
    double datagen_seconds, train_seconds, eval_seconds;
    double first_loss, final_loss, train_gemm_tflops;
    long long steps;
    double class_agreement, gated_fraction, gated_class_agreement;
    double defl_median, defl_p99, defl_p999;        /* rad, gated escaped held-out rays */
    double g_rel_median, g_rel_p99, rhit_rel_median, rhit_rel_p99;
} kt_surrogate_report;

void kt_surrogate_opts_default(kt_surrogate_opts* opts);
/* parameters of KT_NET_WIDE; of either network; 0 for an unknown one */
long long kt_surrogate_param_count(void);
long long kt_surrogate_net_param_count(int net);
/* the network the context holds (KT_NET_WIDE until set): set_net drops its parameters;
 * kt_surrogate_load picks the network from the parameter count, kt_surrogate_train from
 * opts->net, kt_surrogate_init initialises the current one */
int kt_surrogate_set_net(kt_ctx* ctx, int net);
int kt_surrogate_get_net(kt_ctx* ctx);
int kt_surrogate_init(kt_ctx* ctx, unsigned long long seed);
int kt_surrogate_load(kt_ctx* ctx, const float* params, long long n);
int kt_surrogate_save(kt_ctx* ctx, float* params, long long n);
/* generates the data with the integrator, trains on the GPU, judges the held-out set */
int kt_surrogate_train(kt_ctx* ctx, const kt_surrogate_opts* opts, kt_surrogate_report* report);
/* network outputs for normalised inputs (device pointers, n x 5 -> n x 8 floats) */
int kt_surrogate_predict(kt_ctx* ctx, const float* inputs, long long n, float* outputs, void* hip_stream);
/* gradient of the mean loss over n rays (device inputs: n x 5 floats, n classes, n x 5
 * targets) into grads (device, param_count floats); the loss to *loss (host) */
int kt_surrogate_gradients(kt_ctx* ctx, const float* inputs, const int8_t* cls, const float* targets, int n,
                           float* grads, double* loss);
/* one Adam step (beta 0.9, 0.999, eps 1e-8) with device gradients, step t >= 1 */
int kt_surrogate_adam_step(kt_ctx* ctx, const float* grads, double lr, int t);
/* the normalisation's shadow table: critical radius along psi and mean shadow radius */
double kt_surrogate_table_radius(kt_ctx* ctx, double a, double inclination, double psi);
double kt_surrogate_shadow_radius(kt_ctx* ctx, double a, double inclination);

/* What the loaded network covers.  With use_surrogate = 1 every ray of a view outside it
 * goes to the integrator: cameras other than Bardeen planes; a spin, inclination or camera
 * radius outside the ranges; a disk that is disabled, retrograde, starts anywhere but the
 * ISCO or ends anywhere but disk_outer; another escape radius or capture factor.  Per ray,
 * |alpha| or |beta| beyond `extent` mean shadow radii goes there too.  kt_surrogate_train
 * sets the domain to the ranges it draws from and the scene of its options; a network given
 * by kt_surrogate_load or kt_surrogate_init keeps the current one (the default training's
 * until set), so a caller loading parameters trained on another scene sets it here. */
typedef struct kt_surrogate_domain {
    double spin_min, spin_max;                /* 0, 0.998                                  */
    double inclination_min, inclination_max;  /* rad: 5 deg, 85 deg                        */
    double r_cam_min, r_cam_max;              /* 10, 1000                                  */
    double extent;                            /* 1.5 (at most: the inputs' normalisation)  */
    double disk_outer, escape_radius, capture_factor;  /* 20, 1000, 1.01               */
} kt_surrogate_domain;
void kt_surrogate_domain_default(kt_surrogate_domain* domain);
int kt_surrogate_get_domain(kt_ctx* ctx, kt_surrogate_domain* domain);
int kt_surrogate_set_domain(kt_ctx* ctx, const kt_surrogate_domain* domain);

/* ---- HDR colour of traced rays ---------------------------------------------------
 * Replaces the colouring at the end of the simulator's trace (shaders/trace.frag:
 * diskEmission 373-403, shiftedBlackbody 330-335, skyColor/gridSky/imageSky 1419-1458,
 * output 1893-1906) for the traced per-ray outputs: disk rays get the thermal emission of
 * a zero-torque thin disk (T^4 ~ r^-3 (1 - sqrt(r_in / r)), peak = disk_temperature)
 * seen as a blackbody at g T, escaped rays the sky at their asymptotic direction,
 * captured and unfinished rays black.  Output: 4 half floats (RGBA, A = 1) per ray. */
enum { KT_SKY_BLACK = 0, KT_SKY_GRID = 1, KT_SKY_IMAGE = 2 };
typedef struct kt_shade_opts {
    double disk_temperature;   /* K at the emission peak (10000)                       */
    double disk_brightness;    /* multiplies the disk colour (1)                        */
    int sky;                   /* KT_SKY_* (grid)                                       */
    double sky_brightness;     /* multiplies the sky colour (1)                         */
    double grid_line;          /* grid sky: line half-width, in grid cells (0.02)       */
    const float* sky_image;    /* KT_SKY_IMAGE: device, equirectangular RGB floats,     */
    int sky_width, sky_height; /*   row 0 at latitude -90 deg, column 0 at longitude -pi */
} kt_shade_opts;
void kt_shade_opts_default(kt_shade_opts* opts);
/* rays: the device outputs of a kt_trace_views call on the same views (each of the sky
 * angles, disk radius and g in FP64, or else its FP32 field); rgba: device, 4 x uint16
 * (IEEE half) per ray */
int kt_shade_views(kt_ctx* ctx, const kt_view* views, int n_views, const kt_rays_out* rays,
                   const kt_shade_opts* opts, void* rgba, void* hip_stream);
/* n linear-sRGB blackbody colours over log10 T in [2.5, 5.3], luminance 1 at 6500 K
 * (src/blackbody.cpp:48-60) */
int kt_blackbody_table(int n, float* rgb);

/* ---- What the Event Horizon Telescope sees (src/eht.h:100-153) -----------------------
 * Each pixel of a distant observer's Bardeen plane is traced back from r = 1000 M in FP64
 * (the same DP5 stepper as kt_trace_views) and the thermal synchrotron light of a hot
 * accretion flow gathered along it by radiative transfer after every accepted step inside
 * r = 60 M (src/eht.cpp:260-319).  Plasma in cgs, intensities in erg s^-1 cm^-2 Hz^-1
 * sr^-1, fluxes in janskys.  opts: rtol, atol and max_steps of kt_trace_opts are used
 * (NULL = defaults); the path always computes in FP64. */
typedef struct kt_eht_flow {   /* eht::Flow (src/eht.h:104-111)                       */
    double n0;                 /* electrons cm^-3 at r = 1 M (1e5)                      */
    double T0;                 /* electron temperature (K) at r = 1 M (4e11)            */
    double beta;               /* plasma beta (10)                                      */
    double keplerian;          /* rotation as a fraction of Keplerian (0.9)             */
    double height;             /* H/R (1)                                               */
    double sin_theta_b;        /* field angle to the ray, sqrt(2/3) (0.8165)            */
} kt_eht_flow;
typedef struct kt_eht_view {   /* eht::View (src/eht.h:118-125)                       */
    double a;                  /* spin (0.9)                                            */
    double inclination;        /* radians (0.3)                                         */
    double half_width;         /* M (16)                                                */
    int size;                  /* pixels across (128)                                   */
    double frequency;          /* Hz (230e9)                                            */
    double mass_solar;         /* sets the path lengths in cm (6.5e9)                   */
} kt_eht_view;
typedef struct kt_eht_out {    /* eht::Image / eht::Ray, SoA, rows from the top        */
    double* intensity;         /* I_nu at the observer (required)                       */
    uint8_t* shadow;           /* 1 where the ray fell in (optional)                    */
    double* min_radius;        /* optional                                              */
    int32_t* steps;            /* DP5 attempts (optional)                               */
    int on_device;             /* 1: device pointers; 0: host (staged, synchronous)    */
} kt_eht_out;
void kt_eht_flow_default(kt_eht_flow* flow);
void kt_eht_view_default(kt_eht_view* view);
/* eht::render (src/eht.cpp:321-347): view->size^2 pixels */
int kt_eht_render(kt_ctx* ctx, const kt_eht_view* view, const kt_eht_flow* flow, const kt_trace_opts* opts,
                  const kt_eht_out* out, void* hip_stream);
/* eht::traceRay with the light (src/eht.cpp:260-319) for n explicit rays: ab = n (alpha,
 * beta) pairs in M, on the device when out->on_device, else on the host */
int kt_eht_trace_rays(kt_ctx* ctx, const kt_eht_view* view, const kt_eht_flow* flow, const double* ab,
                      long long n, const kt_trace_opts* opts, const kt_eht_out* out, void* hip_stream);
/* eht::flux (src/eht.cpp:349-354): Jy of a size x size device image spanning
 * [-half_width, half_width]^2 whose M subtends theta_guas microarcseconds; synchronous */
int kt_eht_flux(kt_ctx* ctx, const double* intensity, int size, double half_width, double theta_guas,
                double* jy, void* hip_stream);
/* eht::calibrate (src/eht.cpp:432-457): the n0 that gives target_jy, by a secant in
 * log n0 on 48 x 48 renders; synchronous */
int kt_eht_calibrate(kt_ctx* ctx, const kt_eht_view* view, const kt_eht_flow* flow, double target_jy,
                     double theta_guas, const kt_trace_opts* opts, double* n0, void* hip_stream);
/* eht::angularScale (src/eht.cpp:46-48): microarcseconds that M subtends */
double kt_eht_angular_scale(double mass_solar, double distance_pc);

/* ---- Massive particles (debris) ------------------------------------------------------
 * Replaces the geodesic part of the debris update (shaders/particles_update.comp:207-255,
 * particle_common.glsl) and kerr::stepCoordinateTime (src/kerr.cpp:544-561): each moving
 * particle (mu = 1, constants E, L, Q) advances by dt of ingoing Kerr-Schild coordinate
 * time, in FP64 with the DP5 stepper, on the device.  The state lives in the integrator's
 * own variables so that frame after frame it never round-trips through theta:
 * mu = cos(theta), vmu = Sigma dmu/dtau = -sin(theta) Sigma dtheta/dtau, vr = Sigma dr/dtau
 * (kerr::State::vr), phi the Kerr-Schild phi.  A particle plunges below r_stop
 * (<= 0: r+ + 0.02, src/trajectory.cpp:137) or escapes beyond r_escape; its fate says so
 * and later calls skip it. */
enum { KT_PARTICLE_MOVING = 0, KT_PARTICLE_PLUNGED = 1, KT_PARTICLE_ESCAPED = 2, KT_PARTICLE_UNFINISHED = 3 };
typedef struct kt_particles {
    double* r;
    double* mu;
    double* phi;
    double* vr;
    double* vmu;
    const double* E;      /* per particle */
    const double* L;
    const double* Q;
    int8_t* fate;         /* in: KT_PARTICLE_MOVING to advance; out: the fate        */
    int32_t* steps;       /* optional: DP5 attempts of this call                     */
    double* h;            /* optional: the first step size to try (<= 0: min(dt, 0.01 r)); */
                          /*   out: the controller's next step, for the next call    */
    int on_device;        /* 1: device pointers on the call's stream; 0: host (sync) */
} kt_particles;
/* opts: rtol, atol and max_steps (per particle per call) of kt_trace_opts, NULL = defaults */
int kt_particles_advance(kt_ctx* ctx, double a, const kt_particles* particles, long long n, double dt,
                         double r_stop, double r_escape, const kt_trace_opts* opts, void* hip_stream);
/* A particle launched with 3-velocity v (along e_r, e_theta, e_phi) relative to the ZAMO at
 * (r, theta), r outside the horizon: kerr::launchFromZamo with mu = 1 (src/kerr.cpp:365-406).
 * out = {E, L, Q, vr, vmu}. */
int kt_particle_from_zamo(double a, double r, double theta, const double v[3], double out[5]);

/* kerr::observerFrame (src/kerr.cpp:432-456): the orthonormal frame of the observer of
 * 4-velocity u (Kerr-Schild) at (r, theta), rows as kt_view.frame; u is normalised first.
 * KT_EINVAL unless u is future-directed and timelike. */
int kt_observer_frame(double a, double r, double theta, const double u[4], double e[16]);
/* The 4-velocity (ingoing Kerr-Schild, contravariant) of a free-falling massive observer
 * with constants E, L, Q at (r, theta), moving outward in r when radial_sign > 0 and
 * toward larger theta when polar_sign > 0: kerr::fourVelocity (src/kerr.cpp:414-418) with
 * Sigma dr/dtau = radial_sign sqrt(R), Sigma dtheta/dtau = polar_sign sqrt(Theta).  For
 * KT_CAMERA_OBSERVER views (u_obs), e.g. a camera falling from rest at infinity (E = 1,
 * L = Q = 0, radial_sign = -1), inside the horizon too.  Returns KT_EINVAL where R or
 * Theta is negative by more than rounding. */
int kt_geodesic_four_velocity(double a, double E, double L, double Q, double r, double theta, int radial_sign,
                              int polar_sign, double u[4]);

/* Host-side closed forms (src/kerr.cpp:157-206, 296-299).  Past a = 1 (a naked
 * singularity, up to a = 1.5) the horizon radius is 0, the ISCO is the root of the
 * marginal-stability condition (inside r = 1 prograde), and the prograde photon orbit is 0. */
double kt_horizon_radius(double a);
double kt_isco_radius(double a, int prograde);
double kt_photon_orbit_radius(double a, int prograde);
double kt_circular_omega(double a, double r, int prograde);

/* Critical curve of spin a at inclination i: writes up to cap points, returns the
 * count (src/eht.cpp:52-80; none past a = 1); radius of a closed curve along direction psi
 * (:82-97). */
int kt_critical_curve(double a, double inclination, int samples, double* alpha, double* beta, int cap);
double kt_critical_radius(const double* alpha, const double* beta, int n, double psi);

#ifdef __cplusplus
}
#endif
#endif
