// Host-side interface between the C ABI (mpmx_api.cpp) and the templated GPU
// engine (mpmx_step.hip).  No HIP types appear here.
#pragma once
#include <stdint.h>

#include <string>

struct mpmx_grid_desc;

namespace mpmx {

struct EngineBase {
  virtual ~EngineBase() {}
  virtual int dim() const = 0;
  virtual int precision() const = 0;
  virtual void set_particles(int64_t n, const void* X, const void* V, const void* F, const void* B,
                             const void* mass, const void* volume, const void* mu, const void* lambda,
                             const uint8_t* valid) = 0;
  virtual void set_batch_particles(int nscenes, const int64_t* counts, const void* X, const void* V, const void* F,
                                   const void* B, const void* mass, const void* volume, const void* mu,
                                   const void* lambda, const uint8_t* valid) = 0;
  virtual int scene_count() const = 0;
  virtual void get_particles(void* X, void* V, void* F, void* B, uint8_t* valid) = 0;
  virtual int64_t particle_count() const = 0;
  virtual void get_material(void* mass, void* volume, void* mu, void* lambda) = 0;
  virtual void set_time(double t) = 0;
  virtual int plastic_kind() const = 0;
  virtual void get_lame0(void* mu0, void* lambda0) = 0;
  virtual void restore_plastic(const void* Fp, const void* mu0, const void* lambda0, const void* plastic_def,
                               const void* rho_F, const void* cohesion, const void* vc) = 0;
  virtual void set_constitutive(int model, double neo_threshold) = 0;
  virtual void set_gravity(const double* g) = 0;
  virtual void add_collision_box(int type, double friction, const double* lo, const double* hi) = 0;
  virtual void clear_collision_objects() = 0;
  virtual void step(double dt, int nsteps, bool wait) = 0;
  virtual void step_stage(double dt, int stage) = 0;
  virtual void synchronize() = 0;
  virtual double time() const = 0;
  virtual void enable_grid_stats(bool on) = 0;
  virtual void grid_stats(double* mass, double* momentum) = 0;
  virtual void particle_stats(double* mass, double* momentum, int64_t* valid_count) = 0;
  virtual void grid_totals(int which, double* out) = 0;
  virtual void particle_totals(double* out) = 0;
  virtual int64_t grid_node_count() const = 0;
  virtual void get_grid(void* mass, void* velocity) = 0;
  virtual void set_reorder_interval(int interval) = 0;
  virtual void set_rebin_policy(double mover_fraction, int max_interval) = 0;
  virtual void set_launch_config(const mpmx_launch_config& cfg) = 0;
  virtual mpmx_launch_config launch_config() const = 0;
  virtual void autotune(double dt, int trial_steps, mpmx_launch_config* chosen, double* times_ms, int max_times,
                        int* ntimes) = 0;
  virtual void bin_stats(int64_t* rebins, int64_t* movers, int64_t* slow) = 0;
  virtual void grid_v_upper_bound(double* out) = 0;
  virtual void compute_dt(double cfl, double max_dt, double* dt) = 0;
  virtual void advance_frames(int nframes, double frame_dt, double cfl, double min_dt, double max_dt, double* dts,
                              int64_t capacity, int64_t* nsub) = 0;
  virtual void set_plasticity(bool enable, double theta_c, double theta_s, double max_hardening, double xi) = 0;
  virtual void get_plastic_state(void* Fp, void* mu, void* lambda) = 0;
  virtual void set_reflection(bool enable, bool full, double friction, const int* sides) = 0;
  virtual void set_dt_limits(bool strong, bool sound, double cfl, double cfl_F, double cfl_sound, double min_dt) = 0;
  virtual void max_sound_speed(double* out) = 0;
  virtual double last_dt() const = 0;
  virtual void set_drucker_prager(bool enable, double a0, double a1, double a3, double a4, double cohesion,
                                  bool volume_correction) = 0;
  virtual void get_drucker_prager_state(void* plastic_def, void* rho_F, void* cohesion, void* vc) = 0;
  virtual void tape_begin(int checkpoint_every) = 0;
  virtual void backward(const void* dX, const void* dV, const void* dF, const void* dB, void* gX, void* gV, void* gF,
                        void* gB, double* gmu, double* glambda) = 0;
  virtual void enable_timing(int mode) = 0;  // 0 off, 1 every phase, 2 the fused tile kernel only
  virtual void reset_timing() = 0;
  virtual bool phase_time(const char* phase, double* total_ms, int64_t* count) = 0;
  virtual void stress(int64_t n, const void* F, const void* volume, const void* mu, const void* lambda,
                      void* tau) = 0;
  virtual void svd(int64_t n, const void* F, void* U, void* s, void* V) = 0;
  virtual void sort_keys(int64_t n, const uint32_t* keys, int key_bits, uint32_t* keys_out,
                         uint32_t* order_out) = 0;
  virtual void sort_keys_noalloc(int64_t n, const uint32_t* keys, int key_bits, uint32_t* keys_out,
                                 uint32_t* order_out) = 0;
};

// Throws mpmx::Error on failure.
EngineBase* make_engine(const mpmx_grid_desc* descs, int nscenes);

struct Error {
  int code;
  std::string msg;
};

}  // namespace mpmx
