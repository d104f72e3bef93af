// kt_surrogate.hpp -- the learned surrogate (MLP 5-256-256-256-256-8, BF16 MFMA).
//
// Inputs per ray (all in [-1, 1] over the training ranges):
//   log10 r_cam - 2,  (cos i - 0.5415) / 0.4545,  (a - 0.499) / 0.499,
//   alpha / (1.5 R_sh),  beta / (1.5 R_sh)
// with R_sh the mean critical-curve radius of (a, i) from the shadow table.  Outputs:
// three class logits (escaped, captured, disk); the escape direction's components along
// the ray's straight-line frame (e1 the undeflected direction, e2 from the line's closest
// point towards the hole, e3 = e1 x e2); log r_hit and log g.  The input layer
// runs in FP32, layers 1-4 on the matrix cores.  A ray is answered by the surrogate
// when its largest class probability exceeds gate_prob and it lies outside
// |rho / rho_c(psi) - 1| <= gate_ring of the critical curve; all others go to the
// integrator through a device-side queue.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kt_internal.hpp"
#include "kt_nn.hpp"

namespace kts {

struct Surrogate;

struct TrainOpts {
    long long n_train = 1ll << 24;
    long long n_holdout = 1ll << 20;
    int epochs = 192;
    int batch = 65536;
    double lr = 1e-3;           // cosine schedule from lr to 0
    double w_sky = 30.0, w_disk = 1.0;
    double band = 0.02;         // rays within this relative distance of the critical curve ...
    double band_factor = 4.0;   // ... are drawn this many times more often
    uint64_t seed = 0x5EED0004ull;
    int label_precision = 1;    // KT_PREC_* of the integrator labelling the training rays
    double disk_outer = 20.0, escape_radius = 1000.0, capture_factor = 1.01;
    double rtol = 1e-10, atol = 1e-12, rtol32 = 1e-6, atol32 = 1e-7, restep_drift = 1e-6;
    int net = 0;                // ktn::kNet*: the network to train
    double edge_band = 0.06;    // rays hitting the disk within this of its inner edge ...
    int edge_copies = 4;        // ... get this many jittered neighbours in the training set
    int verbose = 0;            // progress lines on stderr
};

// What a trained network covers (the ranges its training rays are drawn from and the scene
// its labels come from).  Views or rays outside it go to the integrator.
struct Domain {
    double spin_min = 0.0, spin_max = 0.998;
    double inc_min = 5.0 * 3.14159265358979323846 / 180.0, inc_max = 85.0 * 3.14159265358979323846 / 180.0;
    double r_cam_min = 10.0, r_cam_max = 1000.0;
    double extent = 1.5;  // |alpha|, |beta| <= extent x the mean shadow radius (at most 1.5)
    double disk_outer = 20.0, escape_radius = 1000.0, capture_factor = 1.01;
};

struct TrainReport {
    double datagen_seconds = 0, train_seconds = 0, eval_seconds = 0;
    double first_loss = 0, final_loss = 0;
    double train_gemm_tflops = 0;       // achieved BF16 MFMA rate over the training steps
    long long steps = 0;
    // held-out rays against FP64 labels
    double class_agreement = 0;         // all held-out rays
    double gated_fraction = 0;          // fraction the gate lets the surrogate answer
    double gated_class_agreement = 0;   // among those
    double defl_median = 0, defl_p99 = 0, defl_p999 = 0;  // rad, gated escaped rays
    double g_rel_median = 0, g_rel_p99 = 0, rhit_rel_median = 0, rhit_rel_p99 = 0;  // gated disk rays
};

Surrogate* create(int device, hipError_t* err);
void destroy(Surrogate* s);
bool ready(const Surrogate* s);

// which network the surrogate holds (ktn::kNetWide or ktn::kNetFast); changing it drops the
// parameters (init_params, set_params or train give new ones)
hipError_t set_net(Surrogate* s, int net);
int get_net(const Surrogate* s);

hipError_t init_params(Surrogate* s, uint64_t seed);
hipError_t set_params(Surrogate* s, const float* host, long long n);
hipError_t get_params(Surrogate* s, float* host, long long n);

// Network on given inputs (device, n x 5 floats) -> outputs (device, n x 8 floats).
hipError_t predict(Surrogate* s, const float* feat, long long n, float* out, hipStream_t stream);
// One backward pass of the mean loss over n rays (device inputs); grads (device, kParams).
hipError_t gradients(Surrogate* s, const float* feat, const int8_t* cls, const float* tgt, int n, float* grads,
                     double* loss, hipStream_t stream);
// One Adam step with the given gradients (device) and learning rate, step t (1-based).
hipError_t adam_step(Surrogate* s, const float* grads, double lr, int t, hipStream_t stream);

hipError_t train(Surrogate* s, const TrainOpts& o, TrainReport& rep);

// the network's domain: set by train() from its options, else the default training's
Domain get_domain(const Surrogate* s);
// a covered view's per-view network inputs (the float forms the kernels use): c[0..2] the
// inputs from r_cam, inclination and spin, c[3] = 1 / (1.5 R_sh(a, inc)), c[4] = inc
void view_constants(const Surrogate* s, double a, double sin_i, double cos_i, double r_cam, float c[6]);
void set_domain(Surrogate* s, const Domain& d);

// The fallback queue the gate fills: ray indices and their count (device).  Calls that
// share a queue must not overlap; a plan owns its own so that plans can run concurrently.
struct Queue {
    long long cap = 0;
    long long* list = nullptr;
    unsigned long long* count = nullptr;  // [0..2] rays queued per region, [3] the network's tile counter
    long long* total = nullptr;           // the same count, as the integrator reads it
};
hipError_t queue_reserve(Queue& q, long long n);
void queue_release(Queue& q);

// Surrogate inference over every ray of the batch, the gate, (closed_form) the closed-form
// values of the rays it answered as escaped or disk hits, then the integrator over the rays
// the gate rejects (re-queued through L.ray_list).  q = nullptr uses the surrogate's
// own queue.  ev[1], ev[2] bracket the integrator kernel; ev = nullptr records no events
// (stream capture), and then q must already hold L.total rays.
hipError_t trace_with_surrogate(Surrogate* s, const kti::TraceLaunch& L, double gate_prob, double gate_ring,
                                int closed_form, Queue* q, hipStream_t stream, hipEvent_t* ev);

// critical radius / shadow radius from the table (host), as the kernels see them
double table_critical_radius(const Surrogate* s, double a, double inc, double psi);
double table_shadow_radius(const Surrogate* s, double a, double inc);

}  // namespace kts
