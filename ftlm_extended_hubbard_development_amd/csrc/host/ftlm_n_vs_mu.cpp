// ftlm_n_vs_mu -- drop-in replacement of the reference executable of the
// same name (src/main.cpp:1459-1839) running the hot path on the GPU engine
// through the C ABI of include/ftlm.h.
//
// Same flags, stdout lines, CSV columns, checkpoint/.meta files and block
// order (N_up, N_down, my, mx) as the reference. Per sector, the exact
// blocks are diagonalised in one batched launch before the blocks are
// visited in order, so records and accumulation keep the reference's order.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <numeric>
#include <sstream>
#include <string>
#include <vector>

#include "../../../include/ftlm.h"
#include "ftlm_host.h"

using namespace ftlmh;

namespace {

void check(int rc, const char* what) {
  if (rc != FTLM_OK) die(std::string(what) + ": " + ftlm_last_error());
}

struct Ctx {
  ftlm_ctx* c = nullptr;
  ~Ctx() { ftlm_destroy(c); }
};

}  // namespace

int main(int argc, char** argv) {
  try {
    Options o;
    if (!parse_options(argc, argv, o, false)) return 0;
    const int sites = o.lx * o.ly;
    const bool multi = o.betas.size() > 1;
    const std::vector<double> mu = linspace(o.mu_min, o.mu_max, o.mu_count);
    const bool trace = !o.trace_csv.empty();
    const bool dbg = o.dbg_nup >= 0 && o.dbg_ndn >= 0 && o.dbg_mx >= 0 && o.dbg_my >= 0;
    if (!o.checkpoint.empty() && (trace || dbg))
      die("--checkpoint is not currently compatible with trace/debug block output.");
    check_or_write_metadata(o);
    ensure_header(o.checkpoint);
    std::uint64_t valid = 0;
    auto ckpt = read_checkpoint(o.checkpoint, &valid);
    truncate_to(o.checkpoint, valid);
    const int threads = requested_threads(o);

    std::cout << "Reference used: "
              << "reference_ftlm_hub_cond/fltm_hub_cond/hubTri2Dcond_omp.f"
              << " and cond_spect_omp.f\n";
    std::cout << "Running FTLM with translation-symmetry k blocks on a " << o.lx << "x" << o.ly
              << " rectangular lattice (" << sites << " sites)\n";
    std::cout << "beta=";
    for (std::size_t i = 0; i < o.betas.size(); ++i) std::cout << (i ? "," : "") << o.betas[i];
    std::cout << " U=" << o.u << " V=" << o.v << " tx=" << o.tx << " ty=" << o.ty << " tp=" << o.tp
              << " phix=" << o.phix << " phiy=" << o.phiy << " threads=" << threads
              << " exact_block_threshold=" << o.threshold << "\n";
    if (!o.checkpoint.empty())
      std::cout << "checkpoint=" << o.checkpoint << " loaded_blocks=" << ckpt.size() << "\n";

    Ctx ctx;
    const ftlm_model model{o.lx, o.ly, o.tx, o.ty, o.tp, o.phix, o.phiy, o.u, o.v};
    check(ftlm_create(&model, o.device, &ctx.c), "GPU engine unavailable");

    std::vector<Spectrum> stored;
    Thermo single(mu.size());
    std::ofstream tout;
    if (trace) {
      tout.open(o.trace_csv);
      if (!tout) die("Failed to open trace CSV: " + o.trace_csv);
      tout << "nup,ndown,mx,my,particles,basis_dim,mu,"
           << "method,z_used,z_ed,z_ratio,n_used_total,n_ed_total,n_total_diff,"
           << "n2_used_total,n2_ed_total,n2_total_diff,"
           << "min_ftlm,min_ed,sum_sample_weights\n";
      tout << std::setprecision(15);
    }
    bool wrote_dbg = false;
    std::ofstream tj;
    if (!o.timing_json.empty()) {
      tj.open(o.timing_json);
      if (!tj) die("Failed to open timing JSON: " + o.timing_json);
      tj << "[\n";
    }
    bool first_tj = true;
    const int L = sites;
    const std::size_t R = (std::size_t)o.samples;

    for (int nup = 0; nup <= L; ++nup) {
      for (int ndn = 0; ndn <= L; ++ndn) {
        std::uint64_t full = 0;
        std::int64_t parents = 0;
        check(ftlm_sector(ctx.c, nup, ndn, &full, &parents), "sector build");
        if (full == 0) continue;
        if (full > o.max_sector_dim) {
          std::ostringstream msg;
          msg << "Sector (Nup=" << nup << ", Ndown=" << ndn << ") has full dimension " << full
              << ", which exceeds --max-sector-dim=" << o.max_sector_dim;
          die(msg.str());
        }
        // block dimensions and the exact batch of this sector
        std::vector<std::int64_t> dims(L);
        std::vector<int> exm, exy;
        std::vector<int> exidx(L, -1);
        for (int my = 0; my < o.ly; ++my)
          for (int mx = 0; mx < o.lx; ++mx) {
            const int k = my * o.lx + mx;
            check(ftlm_block(ctx.c, mx, my, &dims[k]), "block build");
            if (dims[k] == 0) continue;
            if (ckpt.count(BlockKey{nup, ndn, mx, my})) continue;
            const bool use_exact = (std::size_t)dims[k] <= o.threshold;
            const bool dbg_here = dbg && nup == o.dbg_nup && ndn == o.dbg_ndn && mx == o.dbg_mx && my == o.dbg_my;
            if (use_exact || trace || dbg_here) {
              exidx[k] = (int)exm.size();
              exm.push_back(mx);
              exy.push_back(my);
            }
          }
        std::vector<double> exeig;
        std::vector<std::int64_t> exoff(exm.size() + 1, 0);
        if (!exm.empty()) {
          std::int64_t tot = 0;
          for (std::size_t i = 0; i < exm.size(); ++i) tot += dims[exy[i] * o.lx + exm[i]];
          exeig.resize((std::size_t)tot);
          const auto t0 = std::chrono::steady_clock::now();
          check(ftlm_exact_blocks(ctx.c, (int)exm.size(), exm.data(), exy.data(), exeig.data(), exoff.data()),
                "exact diagonalization");
          if (tj) {
            const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            tj << (first_tj ? "" : ",\n") << "{\"nup\":" << nup << ",\"ndn\":" << ndn << ",\"method\":\"exact_batch\","
               << "\"blocks\":" << exm.size() << ",\"ms\":" << ms << "}";
            first_tj = false;
          }
        }
        std::size_t active = 0;
        for (int my = 0; my < o.ly; ++my) {
          for (int mx = 0; mx < o.lx; ++mx) {
            const int k = my * o.lx + mx;
            const std::int64_t dim = dims[k];
            active += (std::size_t)dim;
            if (dim == 0) continue;
            const int particles = nup + ndn;
            const BlockKey key{nup, ndn, mx, my};
            const auto hit = ckpt.find(key);
            if (hit != ckpt.end()) {
              const Spectrum& s = hit->second;
              if (s.dim != (std::uint64_t)dim || s.particles != particles)
                die("Checkpoint block metadata does not match current run.");
              if (multi) stored.push_back(s);
              else
                accumulate(s.eig.data(), s.w.empty() ? nullptr : s.w.data(), s.eig.size(), s.prefactor, s.particles,
                           o.beta, mu, single, threads);
              continue;
            }
            const bool use_exact = (std::size_t)dim <= o.threshold;
            Spectrum spec;
            spec.particles = particles;
            spec.dim = (std::uint64_t)dim;
            spec.exact = use_exact;
            if (!use_exact) {
              const auto t0 = std::chrono::steady_clock::now();
              check(ftlm_block(ctx.c, mx, my, nullptr), "block build");
              const std::size_t m = std::min<std::size_t>((std::size_t)o.steps, (std::size_t)dim);
              spec.eig.resize(R * m);
              spec.w.resize(R * m);
              std::int64_t cnt = 0;
              check(ftlm_ftlm_block(ctx.c, o.samples, o.steps, ftlm_block_seed(o.seed, nup, ndn, mx, my), o.reduction,
                                    o.per_vector ? 1 : 0, spec.eig.data(), spec.w.data(), &cnt),
                    "FTLM block");
              spec.eig.resize((std::size_t)cnt);
              spec.w.resize((std::size_t)cnt);
              spec.prefactor = (double)dim;
              if (tj) {
                const double ms =
                    std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
                tj << (first_tj ? "" : ",\n") << "{\"nup\":" << nup << ",\"ndn\":" << ndn << ",\"mx\":" << mx
                   << ",\"my\":" << my << ",\"dim\":" << dim << ",\"method\":\"ftlm\",\"ms\":" << ms << "}";
                first_tj = false;
              }
            }
            const double* exact_eig = nullptr;
            std::size_t exact_n = 0;
            if (exidx[k] >= 0) {
              exact_eig = exeig.data() + exoff[exidx[k]];
              exact_n = (std::size_t)(exoff[exidx[k] + 1] - exoff[exidx[k]]);
            }
            if (use_exact) {
              spec.eig.assign(exact_eig, exact_eig + exact_n);
              spec.prefactor = 1.0;
            }
            const double sumw = use_exact ? 1.0 : std::accumulate(spec.w.begin(), spec.w.end(), 0.0);
            if (trace) {
              for (double m : mu) {
                const BlockThermo used = use_exact
                                             ? raw_thermo(exact_eig, nullptr, exact_n, 1.0, particles, o.beta, m)
                                             : raw_thermo(spec.eig.data(), spec.w.data(), spec.eig.size(),
                                                          (double)dim, particles, o.beta, m);
                const BlockThermo ex = raw_thermo(exact_eig, nullptr, exact_n, 1.0, particles, o.beta, m);
                const double ratio = ex.z > 0.0 ? used.z / ex.z : 0.0;
                tout << nup << "," << ndn << "," << mx << "," << my << "," << particles << "," << dim << "," << m
                     << "," << (use_exact ? "exact" : "ftlm") << "," << used.z << "," << ex.z << "," << ratio << ","
                     << used.n_total << "," << ex.n_total << "," << (used.n_total - ex.n_total) << ","
                     << used.n2_total << "," << ex.n2_total << "," << (used.n2_total - ex.n2_total) << ","
                     << used.min_shifted << "," << ex.min_shifted << "," << sumw << "\n";
              }
            }
            if (dbg && nup == o.dbg_nup && ndn == o.dbg_ndn && mx == o.dbg_mx && my == o.dbg_my) {
              std::cout << "Selected block debug Nup=" << nup << " Ndown=" << ndn << " mx=" << mx << " my=" << my
                        << " basis_dim=" << dim << " method=" << (use_exact ? "exact" : "ftlm")
                        << " sum_sample_weights=" << sumw << "\n";
              for (double m : mu) {
                const BlockThermo used = use_exact
                                             ? raw_thermo(exact_eig, nullptr, exact_n, 1.0, particles, o.beta, m)
                                             : raw_thermo(spec.eig.data(), spec.w.data(), spec.eig.size(),
                                                          (double)dim, particles, o.beta, m);
                const BlockThermo ex = raw_thermo(exact_eig, nullptr, exact_n, 1.0, particles, o.beta, m);
                const double ratio = ex.z > 0.0 ? used.z / ex.z : 0.0;
                std::cout << "  debug mu=" << std::setw(12) << m << " z_used=" << std::setw(14) << used.z
                          << " z_ed=" << std::setw(14) << ex.z << " ratio=" << std::setw(14) << ratio
                          << " n_diff=" << std::setw(14) << (used.n_total - ex.n_total) << "\n";
              }
              if (!o.debug_csv.empty()) {
                // write_selected_block_debug (main.cpp:1233-1267)
                std::ofstream dout(o.debug_csv);
                if (!dout) die("Failed to open selected block debug file: " + o.debug_csv);
                dout << "nup,ndown,mx,my,basis_dim,mu,z_ftlm,z_ed,z_ratio,"
                     << "n_ftlm_total,n_ed_total,n_diff,n2_ftlm_total,n2_ed_total,n2_diff,"
                     << "min_ftlm,min_ed,sum_sample_weights\n";
                dout << std::setprecision(15);
                const bool have_ftlm = !use_exact;
                const double sw = have_ftlm ? std::accumulate(spec.w.begin(), spec.w.end(), 0.0) : 0.0;
                for (double m : mu) {
                  const BlockThermo f = have_ftlm ? raw_thermo(spec.eig.data(), spec.w.data(), spec.eig.size(),
                                                               (double)dim, particles, o.beta, m)
                                                  : BlockThermo{};
                  const BlockThermo ex = raw_thermo(exact_eig, nullptr, exact_n, 1.0, particles, o.beta, m);
                  const double ratio = ex.z > 0.0 ? f.z / ex.z : 0.0;
                  dout << nup << "," << ndn << "," << mx << "," << my << "," << dim << "," << m << "," << f.z << ","
                       << ex.z << "," << ratio << "," << f.n_total << "," << ex.n_total << ","
                       << (f.n_total - ex.n_total) << "," << f.n2_total << "," << ex.n2_total << ","
                       << (f.n2_total - ex.n2_total) << "," << f.min_shifted << "," << ex.min_shifted << "," << sw
                       << "\n";
                }
                wrote_dbg = true;
              }
            }
            if (multi || !o.checkpoint.empty()) {
              append_record(o.checkpoint, key, spec);
              if (multi) {
                stored.push_back(std::move(spec));
              } else {
                accumulate(spec.eig.data(), spec.w.empty() ? nullptr : spec.w.data(), spec.eig.size(), spec.prefactor,
                           particles, o.beta, mu, single, threads);
              }
            } else {
              accumulate(spec.eig.data(), use_exact ? nullptr : spec.w.data(), spec.eig.size(), spec.prefactor,
                         particles, o.beta, mu, single, threads);
            }
          }
        }
        std::cout << "  sector Nup=" << nup << " Ndown=" << ndn << " full_dim=" << full << " parents=" << parents
                  << " active_k_dim_sum=" << active << "\n";
      }
    }
    if (tj) tj << "\n]\n";
    if (dbg && !o.debug_csv.empty() && !wrote_dbg)
      die("Selected block debug target was not found among active FTLM blocks.");

    if (multi) {
      std::vector<Results> all;
      for (double beta : o.betas) {
        Thermo t(mu.size());
        for (const Spectrum& s : stored)
          accumulate(s.eig.data(), s.w.empty() ? nullptr : s.w.data(), s.eig.size(), s.prefactor, s.particles, beta,
                     mu, t, threads);
        Results r = finalize(t, beta, sites);
        for (std::size_t i = 0; i < mu.size(); ++i)
          std::cout << "beta=" << std::setw(12) << beta << "  mu=" << std::setw(12) << mu[i] << "  n=" << std::setw(12)
                    << r.n[i] << "  kappa=" << std::setw(12) << r.kappa[i] << "\n";
        all.push_back(std::move(r));
      }
      write_multi_beta(o.output, o.betas, mu, all);
    } else {
      const Results r = finalize(single, o.beta, sites);
      for (std::size_t i = 0; i < mu.size(); ++i)
        std::cout << "beta=" << std::setw(12) << o.beta << "  mu=" << std::setw(12) << mu[i] << "  n=" << std::setw(12)
                  << r.n[i] << "  kappa=" << std::setw(12) << r.kappa[i] << "\n";
      write_results(o.output, mu, r);
    }
    std::cout << "Wrote " << o.output << "\n";
    return 0;
  } catch (const std::exception& ex) {
    std::cerr << "Error: " << ex.what() << "\n";
    return 1;
  }
}
