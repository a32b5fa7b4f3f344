// ftlm_ed_n_vs_mu -- drop-in replacement of the reference's exact
// diagonalisation executable (src/ed_main.cpp:49-146): every (N_up, N_down, k)
// block is diagonalised by the GPU engine's batched Hermitian eigensolver and
// accumulated in the reference's block order at the single beta of --beta
// (ed_main.cpp:99). Blocks of all sectors go into the engine's exact queue
// (one launch solves every queued block, so the large blocks of different
// sectors run side by side); sectors are accumulated and reported in order
// when the queue is run: when it is full and at the end.
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "../../../include/ftlm.h"
#include "ftlm_host.h"

using namespace ftlmh;

int main(int argc, char** argv) {
  ftlm_ctx* ctx = nullptr;
  try {
    Options o;
    if (!parse_options(argc, argv, o, true)) return 0;
    const int L = o.lx * o.ly;
    const std::vector<double> mu = linspace(o.mu_min, o.mu_max, o.mu_count);
    std::cout << "Exact diagonalization using the FTLM sector and k-block builder\n";
    std::cout << "Running ED on a " << o.lx << "x" << o.ly << " rectangular lattice (" << L << " sites)\n";
    std::cout << "beta=" << o.beta << " U=" << o.u << " V=" << o.v << " tx=" << o.tx << " ty=" << o.ty
              << " tp=" << o.tp << " phix=" << o.phix << " phiy=" << o.phiy << "\n";
    const ftlm_model model{o.lx, o.ly, o.tx, o.ty, o.tp, o.phix, o.phiy, o.u, o.v};
    if (ftlm_create(&model, o.device, &ctx) != FTLM_OK)
      die(std::string("GPU engine unavailable: ") + ftlm_last_error());
    const int threads = requested_threads(o);
    Thermo t(mu.size());
    auto check = [&](int rc, const char* what) {
      if (rc != FTLM_OK) die(std::string(what) + ": " + ftlm_last_error());
    };
    // sectors whose blocks are queued, in the reference's order; a sector
    // whose first blocks were solved by a run triggered in the middle of it
    // keeps their spectra on the host (solved, in block order)
    struct Pending {
      int nup, ndn;
      std::uint64_t full;
      std::int64_t parents;
      std::size_t active;
      std::vector<std::vector<double>> solved;
      std::vector<std::int64_t> idx;
    };
    std::vector<Pending> pending;
    auto fetch = [&](std::int64_t q) {
      std::int64_t cnt = 0;
      check(ftlm_exact_queue_result(ctx, q, nullptr, &cnt), "exact diagonalization");
      std::vector<double> e((std::size_t)cnt);
      check(ftlm_exact_queue_result(ctx, q, e.data(), &cnt), "exact diagonalization");
      return e;
    };
    // run the queue; accumulate and report every pending sector in order;
    // the sector in progress (cur) moves its queued blocks' spectra to the host
    auto flush = [&](Pending* cur) {
      check(ftlm_exact_queue_run(ctx), "exact diagonalization");
      for (Pending& s : pending) {
        for (const std::vector<double>& e : s.solved)
          accumulate(e.data(), nullptr, e.size(), 1.0, s.nup + s.ndn, o.beta, mu, t, threads);
        for (std::int64_t q : s.idx) {
          const std::vector<double> e = fetch(q);
          accumulate(e.data(), nullptr, e.size(), 1.0, s.nup + s.ndn, o.beta, mu, t, threads);
        }
        std::cout << "  sector Nup=" << s.nup << " Ndown=" << s.ndn << " full_dim=" << s.full
                  << " parents=" << s.parents << " active_k_dim_sum=" << s.active << "\n";
      }
      pending.clear();
      if (cur) {
        for (std::int64_t q : cur->idx) cur->solved.push_back(fetch(q));
        cur->idx.clear();
      }
      check(ftlm_exact_queue_begin(ctx, 0), "exact diagonalization");
    };
    check(ftlm_exact_queue_begin(ctx, 0), "exact diagonalization");
    for (int nup = 0; nup <= L; ++nup) {
      for (int ndn = 0; ndn <= L; ++ndn) {
        std::uint64_t full = 0;
        std::int64_t parents = 0;
        check(ftlm_sector(ctx, nup, ndn, &full, &parents), "sector build");
        if (full == 0) continue;
        if (full > o.max_sector_dim) {
          flush(nullptr);  // the sectors before this one are reported first, as the reference does
          std::ostringstream msg;
          msg << "Sector (Nup=" << nup << ", Ndown=" << ndn << ") has full dimension " << full
              << ", which exceeds --max-sector-dim=" << o.max_sector_dim;
          die(msg.str());
        }
        Pending cur{nup, ndn, full, parents, 0, {}, {}};
        for (int my = 0; my < o.ly; ++my)
          for (int mx = 0; mx < o.lx; ++mx) {
            std::int64_t d = 0;
            check(ftlm_block(ctx, mx, my, &d), "block build");
            cur.active += (std::size_t)d;
            if (d == 0) continue;
            std::int64_t q = 0;
            int rc = ftlm_exact_queue_add(ctx, mx, my, &q);
            if (rc == FTLM_PACK_FULL) {
              flush(&cur);
              rc = ftlm_exact_queue_add(ctx, mx, my, &q);
              if (rc == FTLM_PACK_FULL) die("exact diagonalization: a block does not fit the exact-queue budget");
            }
            check(rc, "exact diagonalization");
            cur.idx.push_back(q);
          }
        pending.push_back(std::move(cur));
      }
    }
    flush(nullptr);
    const Results r = finalize(t, o.beta, L);
    for (std::size_t i = 0; i < mu.size(); ++i)
      std::cout << "mu=" << std::setw(12) << mu[i] << "  n=" << std::setw(12) << r.n[i] << "  kappa=" << std::setw(12)
                << r.kappa[i] << "\n";
    write_results(o.output, mu, r);
    std::cout << "Wrote " << o.output << "\n";
    ftlm_destroy(ctx);
    return 0;
  } catch (const std::exception& ex) {
    ftlm_destroy(ctx);
    std::cerr << "Error: " << ex.what() << "\n";
    return 1;
  }
}
