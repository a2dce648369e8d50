// C ABI of the MI355X MPM step (include/mpmx.h).  Every entry point converts
// exceptions from the engine into a status code and a thread-local message.
#include "mpmx.h"

#include <exception>
#include <new>
#include <string>

#include "mpmx_engine.h"
#include "mpmx_io.h"

static_assert(sizeof(mpmx_totals) == 12 * sizeof(double), "mpmx_totals is twelve packed doubles");

#include <unistd.h>

#include <cstdio>
#include <vector>

struct mpmx_ctx {
  mpmx::EngineBase* engine;
};

namespace {
thread_local std::string g_last_error;

template<class F> int guard(F&& f) {
  try {
    f();
    g_last_error.clear();
    return MPMX_OK;
  } catch (const mpmx::Error& e) {
    g_last_error = e.msg;
    return e.code;
  } catch (const std::bad_alloc&) {
    g_last_error = "host allocation failed";
    return MPMX_EINTERNAL;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return MPMX_EINTERNAL;
  } catch (...) {
    g_last_error = "unknown failure";
    return MPMX_EINTERNAL;
  }
}

int bad(const char* msg) {
  g_last_error = msg;
  return MPMX_EINVAL;
}
}  // namespace

extern "C" {

const char* mpmx_last_error(void) { return g_last_error.c_str(); }

int mpmx_create(const mpmx_grid_desc* grid, mpmx_ctx** out) {
  if (!grid || !out) return bad("null argument");
  *out = nullptr;
  return guard([&] {
    mpmx::EngineBase* e = mpmx::make_engine(grid, 1);
    *out = new mpmx_ctx{e};
  });
}

int mpmx_create_batch(const mpmx_grid_desc* grids, int nscenes, mpmx_ctx** out) {
  if (!grids || !out) return bad("null argument");
  *out = nullptr;
  return guard([&] {
    mpmx::EngineBase* e = mpmx::make_engine(grids, nscenes);
    *out = new mpmx_ctx{e};
  });
}

int mpmx_scene_count(const mpmx_ctx* ctx) { return ctx ? ctx->engine->scene_count() : MPMX_EINVAL; }

int mpmx_set_batch_particles(mpmx_ctx* ctx, int nscenes, const int64_t* counts, const void* X, const void* V,
                             const void* F, const void* B, const void* mass, const void* volume, const void* mu,
                             const void* lambda, const uint8_t* valid) {
  if (!ctx) return bad("null context");
  return guard([&] { ctx->engine->set_batch_particles(nscenes, counts, X, V, F, B, mass, volume, mu, lambda, valid); });
}

int mpmx_destroy(mpmx_ctx* ctx) {
  if (!ctx) return MPMX_OK;
  int rc = guard([&] { delete ctx->engine; });
  delete ctx;
  return rc;
}

int mpmx_set_particles(mpmx_ctx* ctx, int64_t n, const void* X, const void* V, const void* F, const void* B,
                       const void* mass, const void* volume, const void* mu, const void* lambda,
                       const uint8_t* valid) {
  if (!ctx) return bad("null context");
  return guard([&] { ctx->engine->set_particles(n, X, V, F, B, mass, volume, mu, lambda, valid); });
}

int mpmx_get_particles(mpmx_ctx* ctx, void* X, void* V, void* F, void* B, uint8_t* valid) {
  if (!ctx) return bad("null context");
  return guard([&] { ctx->engine->get_particles(X, V, F, B, valid); });
}

int mpmx_get_material(mpmx_ctx* ctx, void* mass, void* volume, void* mu, void* lambda) {
  if (!ctx) return bad("null context");
  return guard([&] { ctx->engine->get_material(mass, volume, mu, lambda); });
}

int64_t mpmx_particle_count(const mpmx_ctx* ctx) { return ctx ? ctx->engine->particle_count() : -1; }

int mpmx_set_constitutive(mpmx_ctx* ctx, int model, double neo_threshold) {
  if (!ctx) return bad("null context");
  return guard([&] { ctx->engine->set_constitutive(model, neo_threshold); });
}

int mpmx_set_gravity(mpmx_ctx* ctx, const double* g) {
  if (!ctx) return bad("null context");
  return guard([&] { ctx->engine->set_gravity(g); });
}

int mpmx_add_collision_box(mpmx_ctx* ctx, int type, double friction, const double* lo, const double* hi) {
  if (!ctx) return bad("null context");
  return guard([&] { ctx->engine->add_collision_box(type, friction, lo, hi); });
}

int mpmx_clear_collision_objects(mpmx_ctx* ctx) {
  if (!ctx) return bad("null context");
  return guard([&] { ctx->engine->clear_collision_objects(); });
}

int mpmx_step(mpmx_ctx* ctx, double dt, int nsteps) {
  if (!ctx) return bad("null context");
  return guard([&] { ctx->engine->step(dt, nsteps, true); });
}

int mpmx_step_async(mpmx_ctx* ctx, double dt, int nsteps) {
  if (!ctx) return bad("null context");
  return guard([&] { ctx->engine->step(dt, nsteps, false); });
}

int mpmx_synchronize(mpmx_ctx* ctx) {
  if (!ctx) return bad("null context");
  return guard([&] { ctx->engine->synchronize(); });
}

double mpmx_time(const mpmx_ctx* ctx) { return ctx ? ctx->engine->time() : 0.0; }

int mpmx_enable_grid_stats(mpmx_ctx* ctx, int enable) {
  if (!ctx) return bad("null context");
  return guard([&] { ctx->engine->enable_grid_stats(enable != 0); });
}

int mpmx_grid_stats(mpmx_ctx* ctx, double* mass, double* momentum) {
  if (!ctx) return bad("null context");
  return guard([&] { ctx->engine->grid_stats(mass, momentum); });
}

int mpmx_particle_stats(mpmx_ctx* ctx, double* mass, double* momentum, int64_t* valid_count) {
  if (!ctx) return bad("null context");
  return guard([&] { ctx->engine->particle_stats(mass, momentum, valid_count); });
}

int mpmx_step_stage(mpmx_ctx* ctx, double dt, int stage) {
  if (!ctx) return bad("null context");
  return guard([&] { ctx->engine->step_stage(dt, stage); });
}

const char* mpmx_stage_name(int stage) {
  static const char* const names[MPMX_STAGES] = {"simulated-particles", "update-weights", "prepare-scatter", "p2g",
                                                 "forces", "reduce-dt", "g2p", "plasticity-hardening"};
  return stage >= 0 && stage < MPMX_STAGES ? names[stage] : nullptr;
}

int mpmx_particle_totals(mpmx_ctx* ctx, mpmx_totals* out) {
  if (!ctx) return bad("null context");
  if (!out) return bad("null output");
  return guard([&] { ctx->engine->particle_totals(&out->mass); });
}

int mpmx_grid_totals(mpmx_ctx* ctx, int which, mpmx_totals* out) {
  if (!ctx) return bad("null context");
  if (!out) return bad("null output");
  return guard([&] { ctx->engine->grid_totals(which, &out->mass); });
}

int64_t mpmx_grid_node_count(const mpmx_ctx* ctx) { return ctx ? ctx->engine->grid_node_count() : -1; }

int mpmx_get_grid(mpmx_ctx* ctx, void* mass, void* velocity) {
  if (!ctx) return bad("null context");
  return guard([&] { ctx->engine->get_grid(mass, velocity); });
}

int mpmx_set_reorder_interval(mpmx_ctx* ctx, int interval) {
  if (!ctx) return bad("null context");
  return guard([&] { ctx->engine->set_reorder_interval(interval); });
}

int mpmx_set_rebin_policy(mpmx_ctx* ctx, double mover_fraction, int max_interval) {
  if (!ctx) return bad("null context");
  return guard([&] { ctx->engine->set_rebin_policy(mover_fraction, max_interval); });
}

int mpmx_set_launch_config(mpmx_ctx* ctx, const mpmx_launch_config* cfg) {
  if (!ctx || !cfg) return bad("null argument");
  return guard([&] { ctx->engine->set_launch_config(*cfg); });
}

int mpmx_get_launch_config(mpmx_ctx* ctx, mpmx_launch_config* cfg) {
  if (!ctx || !cfg) return bad("null argument");
  return guard([&] { *cfg = ctx->engine->launch_config(); });
}

int mpmx_autotune(mpmx_ctx* ctx, double dt, int trial_steps, mpmx_launch_config* chosen, double* times_ms,
                  int max_times, int* ntimes) {
  if (!ctx) return bad("null context");
  if (max_times > 0 && !times_ms) return bad("null times array");
  return guard([&] { ctx->engine->autotune(dt, trial_steps, chosen, times_ms, max_times, ntimes); });
}

int mpmx_bin_stats(mpmx_ctx* ctx, int64_t* rebins, int64_t* movers, int64_t* slow) {
  if (!ctx) return bad("null context");
  return guard([&] { ctx->engine->bin_stats(rebins, movers, slow); });
}

int mpmx_grid_v_upper_bound(mpmx_ctx* ctx, double* bound) {
  if (!ctx) return bad("null context");
  if (!bound) return bad("null array");
  return guard([&] { ctx->engine->grid_v_upper_bound(bound); });
}

int mpmx_compute_dt(mpmx_ctx* ctx, double cfl, double max_dt, double* dt) {
  if (!ctx) return bad("null context");
  if (!dt) return bad("null array");
  return guard([&] { ctx->engine->compute_dt(cfl, max_dt, dt); });
}

int mpmx_advance_frames(mpmx_ctx* ctx, int nframes, double frame_dt, double cfl, double min_dt, double max_dt,
                        double* dts, int64_t dts_capacity, int64_t* nsubsteps) {
  if (!ctx) return bad("null context");
  if (dts_capacity < 0 || (dts_capacity > 0 && !dts)) return bad("bad substep buffer");
  return guard([&] { ctx->engine->advance_frames(nframes, frame_dt, cfl, min_dt, max_dt, dts, dts_capacity, nsubsteps); });
}

int mpmx_set_plasticity(mpmx_ctx* ctx, int enable, double theta_c, double theta_s, double max_hardening,
                        double hardening_factor) {
  if (!ctx) return bad("null context");
  return guard([&] { ctx->engine->set_plasticity(enable != 0, theta_c, theta_s, max_hardening, hardening_factor); });
}

int mpmx_get_plastic_state(mpmx_ctx* ctx, void* Fp, void* mu, void* lambda) {
  if (!ctx) return bad("null context");
  return guard([&] { ctx->engine->get_plastic_state(Fp, mu, lambda); });
}

int mpmx_set_reflection(mpmx_ctx* ctx, int enable, int full, double friction, const int* side_types) {
  if (!ctx) return bad("null context");
  return guard([&] { ctx->engine->set_reflection(enable != 0, full != 0, friction, side_types); });
}

int mpmx_set_dt_limits(mpmx_ctx* ctx, int strong_cfl, int sound_cfl, double cfl, double cfl_F, double cfl_sound,
                       double min_dt) {
  if (!ctx) return bad("null context");
  return guard([&] { ctx->engine->set_dt_limits(strong_cfl != 0, sound_cfl != 0, cfl, cfl_F, cfl_sound, min_dt); });
}

int mpmx_max_sound_speed(mpmx_ctx* ctx, double* out) {
  if (!ctx || !out) return bad("null argument");
  return guard([&] { ctx->engine->max_sound_speed(out); });
}

int mpmx_last_dt(mpmx_ctx* ctx, double* out) {
  if (!ctx || !out) return bad("null argument");
  return guard([&] { *out = ctx->engine->last_dt(); });
}

int mpmx_write_particles_file(const char* path, int dim, int precision, int64_t n, const mpmx_particle_array* arrays,
                              int count) {
  return guard([&] { mpmx::io::write_particles(path, dim, precision, n, arrays, count); });
}

int mpmx_read_particles_file(const char* path, int dim, int precision, int64_t* n, mpmx_particle_array* arrays,
                             int count, int* found) {
  if (!n) return bad("null count");
  return guard([&] { *n = mpmx::io::read_particles(path, dim, precision, arrays, count, found); });
}

int mpmx_write_restart_file(const char* path, int precision, double time) {
  if (!path) return bad("null path");
  if (precision != MPMX_FP32 && precision != MPMX_FP64) return bad("precision must be FP32 or FP64");
  return guard([&] { mpmx::io::write_restart(path, precision, time); });
}

int mpmx_read_restart_file(const char* path, double* time) {
  if (!path || !time) return bad("null argument");
  return guard([&] { *time = mpmx::io::read_restart(path); });
}

int mpmx_write_output_files(mpmx_ctx* ctx, const char* frame_dir) {
  if (!ctx || !frame_dir) return bad("null argument");
  return guard([&] {
    mpmx::EngineBase& e = *ctx->engine;
    if (e.scene_count() > 1) throw mpmx::Error{MPMX_ESTATE, "output files take a single-scene context"};
    const int d = e.dim(), prec = e.precision(), ts = prec == MPMX_FP64 ? 8 : 4;
    const int64_t n = e.particle_count();
    const size_t nn = (size_t)n;
    std::vector<char> X(nn * d * ts), V(nn * d * ts), F(nn * d * d * ts), B(nn * d * d * ts), color(nn * 3 * ts, 0);
    std::vector<char> mass(nn * ts), vol(nn * ts), mu(nn * ts), la(nn * ts);
    std::vector<uint8_t> valid(nn);
    e.get_particles(X.data(), V.data(), F.data(), B.data(), valid.data());
    e.get_material(mass.data(), vol.data(), mu.data(), la.data());
    std::vector<mpmx_particle_array> a = {
        {"X", MPMX_ARRAY_VECTOR, X.data()},        {"V", MPMX_ARRAY_VECTOR, V.data()},
        {"F", MPMX_ARRAY_MATRIX, F.data()},        {"B", MPMX_ARRAY_MATRIX, B.data()},
        {"color", MPMX_ARRAY_VECTOR3, color.data()}, {"mass", MPMX_ARRAY_SCALAR, mass.data()},
        {"volume", MPMX_ARRAY_SCALAR, vol.data()}, {"valid", MPMX_ARRAY_BOOL, valid.data()},
        {"mu", MPMX_ARRAY_SCALAR, mu.data()},      {"lambda", MPMX_ARRAY_SCALAR, la.data()}};
    std::vector<char> Fp, mu0, la0, pdef, rho, coh, vc;
    const int kind = e.plastic_kind();
    if (kind) {  // Store_Fp with Store_Lame stores mu0 and lambda0 too (MPM_PARTICLES.cpp:35-43, :95-106)
      Fp.resize(nn * d * d * ts);
      mu0.resize(nn * ts);
      la0.resize(nn * ts);
      e.get_plastic_state(Fp.data(), nullptr, nullptr);
      e.get_lame0(mu0.data(), la0.data());
      a.push_back({"Fp", MPMX_ARRAY_MATRIX, Fp.data()});
      a.push_back({"mu0", MPMX_ARRAY_SCALAR, mu0.data()});
      a.push_back({"lambda0", MPMX_ARRAY_SCALAR, la0.data()});
    }
    if (kind == 2) {  // MPM_DRUCKER_PRAGER's arrays (MPM_DRUCKER_PRAGER.cpp:20-23)
      pdef.resize(nn * ts);
      rho.resize(nn * ts);
      coh.resize(nn * ts);
      vc.resize(nn * ts);
      e.get_drucker_prager_state(pdef.data(), rho.data(), coh.data(), vc.data());
      a.push_back({"plastic_deformation", MPMX_ARRAY_SCALAR, pdef.data()});
      a.push_back({"dp_rho_f", MPMX_ARRAY_SCALAR, rho.data()});
      a.push_back({"dp_cohesion", MPMX_ARRAY_SCALAR, coh.data()});
      a.push_back({"vcp", MPMX_ARRAY_SCALAR, vc.data()});
    }
    const std::string dir = frame_dir;
    mpmx::io::write_particles((dir + "/deformable_object_particles.gz").c_str(), d, prec, n, a.data(), (int)a.size());
    // MPM_EXAMPLE.cpp:64: mpm_particles.gz -> ./deformable_object_particles.gz
    const std::string link = dir + "/mpm_particles.gz";
    std::remove(link.c_str());
    if (symlink("./deformable_object_particles.gz", link.c_str()) != 0)
      throw mpmx::Error{MPMX_EINTERNAL, "cannot create " + link};
    mpmx::io::write_restart((dir + "/restart_data.gz").c_str(), prec, e.time());
  });
}

int mpmx_read_output_files(mpmx_ctx* ctx, const char* frame_dir) {
  if (!ctx || !frame_dir) return bad("null argument");
  return guard([&] {
    mpmx::EngineBase& e = *ctx->engine;
    if (e.scene_count() > 1) throw mpmx::Error{MPMX_ESTATE, "output files take a single-scene context"};
    const int d = e.dim(), prec = e.precision(), ts = prec == MPMX_FP64 ? 8 : 4;
    const std::string dir = frame_dir;
    auto exists = [](const std::string& p) { return access(p.c_str(), R_OK) == 0; };
    // Read_From_File tries the name with ".gz" first (FILE_UTILITIES.cpp:219-235)
    std::string path;
    for (const char* base : {"/mpm_particles", "/deformable_object_particles"})
      for (const char* suf : {".gz", ""})
        if (path.empty() && exists(dir + base + suf)) path = dir + base + suf;
    if (path.empty()) throw mpmx::Error{MPMX_EINVAL, "no particle file in " + dir};
    std::string rpath = exists(dir + "/restart_data.gz") ? dir + "/restart_data.gz" : dir + "/restart_data";
    if (!exists(rpath)) throw mpmx::Error{MPMX_EINVAL, "no restart_data in " + dir};
    const int64_t n = mpmx::io::read_particles(path.c_str(), d, prec, nullptr, 0, nullptr);
    const size_t nn = (size_t)n;
    std::vector<char> X(nn * d * ts), V(nn * d * ts), F(nn * d * d * ts), B(nn * d * d * ts, 0);
    std::vector<char> mass(nn * ts), vol(nn * ts), mu(nn * ts), la(nn * ts);
    std::vector<uint8_t> valid(nn);
    std::vector<char> Fp(nn * d * d * ts), mu0(nn * ts), la0(nn * ts), pdef(nn * ts), rho(nn * ts), coh(nn * ts),
        vc(nn * ts);
    std::vector<mpmx_particle_array> a = {
        {"X", MPMX_ARRAY_VECTOR, X.data()},          {"V", MPMX_ARRAY_VECTOR, V.data()},
        {"F", MPMX_ARRAY_MATRIX, F.data()},          {"mass", MPMX_ARRAY_SCALAR, mass.data()},
        {"volume", MPMX_ARRAY_SCALAR, vol.data()},   {"valid", MPMX_ARRAY_BOOL, valid.data()},
        {"mu", MPMX_ARRAY_SCALAR, mu.data()},        {"lambda", MPMX_ARRAY_SCALAR, la.data()},
        {"B", MPMX_ARRAY_MATRIX, B.data()},          {"Fp", MPMX_ARRAY_MATRIX, Fp.data()},
        {"mu0", MPMX_ARRAY_SCALAR, mu0.data()},      {"lambda0", MPMX_ARRAY_SCALAR, la0.data()},
        {"plastic_deformation", MPMX_ARRAY_SCALAR, pdef.data()}, {"dp_rho_f", MPMX_ARRAY_SCALAR, rho.data()},
        {"dp_cohesion", MPMX_ARRAY_SCALAR, coh.data()}, {"vcp", MPMX_ARRAY_SCALAR, vc.data()}};
    std::vector<int> found(a.size());
    mpmx::io::read_particles(path.c_str(), d, prec, a.data(), (int)a.size(), found.data());
    for (int i = 0; i < 8; i++)
      if (!found[i]) throw mpmx::Error{MPMX_EINVAL, std::string("particle file lacks ") + a[i].name};
    const double t = mpmx::io::read_restart(rpath.c_str());
    e.set_particles(n, X.data(), V.data(), F.data(), B.data(), mass.data(), vol.data(), mu.data(), la.data(),
                    valid.data());
    if (e.plastic_kind())
      e.restore_plastic(found[9] ? Fp.data() : nullptr, found[10] ? mu0.data() : nullptr,
                        found[11] ? la0.data() : nullptr, found[12] ? pdef.data() : nullptr,
                        found[13] ? rho.data() : nullptr, found[14] ? coh.data() : nullptr,
                        found[15] ? vc.data() : nullptr);
    e.set_time(t);
  });
}

int mpmx_set_drucker_prager(mpmx_ctx* ctx, int enable, double a0, double a1, double a3, double a4, double cohesion,
                            int volume_correction) {
  if (!ctx) return bad("null context");
  return guard([&] { ctx->engine->set_drucker_prager(enable != 0, a0, a1, a3, a4, cohesion, volume_correction != 0); });
}

int mpmx_get_drucker_prager_state(mpmx_ctx* ctx, void* plastic_def, void* rho_F, void* cohesion, void* vc) {
  if (!ctx) return bad("null context");
  return guard([&] { ctx->engine->get_drucker_prager_state(plastic_def, rho_F, cohesion, vc); });
}

int mpmx_tape_begin(mpmx_ctx* ctx, int checkpoint_every) {
  if (!ctx) return bad("null context");
  return guard([&] { ctx->engine->tape_begin(checkpoint_every); });
}

int mpmx_backward(mpmx_ctx* ctx, const void* dX, const void* dV, const void* dF, const void* dB, void* gX, void* gV,
                  void* gF, void* gB, double* gmu, double* glambda) {
  if (!ctx) return bad("null context");
  return guard([&] { ctx->engine->backward(dX, dV, dF, dB, gX, gV, gF, gB, gmu, glambda); });
}

int mpmx_enable_timing(mpmx_ctx* ctx, int enable) {
  if (!ctx) return bad("null context");
  if (enable < 0 || enable > 2) return MPMX_EINVAL;
  return guard([&] { ctx->engine->enable_timing(enable); });
}

int mpmx_reset_timing(mpmx_ctx* ctx) {
  if (!ctx) return bad("null context");
  return guard([&] { ctx->engine->reset_timing(); });
}

int mpmx_phase_time(mpmx_ctx* ctx, const char* phase, double* total_ms, int64_t* count) {
  if (!ctx) return bad("null context");
  bool found = false;
  int rc = guard([&] { found = ctx->engine->phase_time(phase, total_ms, count); });
  if (rc == MPMX_OK && !found) return bad("unknown phase");
  return rc;
}

int mpmx_stress(mpmx_ctx* ctx, int64_t n, const void* F, const void* volume, const void* mu, const void* lambda,
                void* tau) {
  if (!ctx) return bad("null context");
  if (n > 0 && (!F || !volume || !mu || !lambda || !tau)) return bad("null array");
  return guard([&] { ctx->engine->stress(n, F, volume, mu, lambda, tau); });
}

int mpmx_svd(mpmx_ctx* ctx, int64_t n, const void* F, void* U, void* s, void* V) {
  if (!ctx) return bad("null context");
  if (n > 0 && (!F || !U || !s || !V)) return bad("null array");
  return guard([&] { ctx->engine->svd(n, F, U, s, V); });
}

int mpmx_sort_keys(mpmx_ctx* ctx, int64_t n, const uint32_t* keys, int key_bits, uint32_t* keys_out,
                   uint32_t* order_out) {
  if (!ctx) return bad("null context");
  if (n > 0 && (!keys || !keys_out || !order_out)) return bad("null array");
  return guard([&] { ctx->engine->sort_keys(n, keys, key_bits, keys_out, order_out); });
}

int mpmx_sort_keys_noalloc(mpmx_ctx* ctx, int64_t n, const uint32_t* keys, int key_bits, uint32_t* keys_out,
                           uint32_t* order_out) {
  if (!ctx) return bad("null context");
  if (n > 0 && (!keys || !keys_out || !order_out)) return bad("null array");
  return guard([&] { ctx->engine->sort_keys_noalloc(n, keys, key_bits, keys_out, order_out); });
}

}  // extern "C"
