// ftlm_ed_n_vs_mu -- drop-in replacement of the reference's exact
// diagonalisation executable (src/ed_main.cpp:49-146): every (N_up, N_down, k)
// block is diagonalised by the GPU engine's batched Hermitian eigensolver
// (one launch per sector) and accumulated in the reference's block order at
// the single beta of --beta (ed_main.cpp:99).
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
    for (int nup = 0; nup <= L; ++nup) {
      for (int ndn = 0; ndn <= L; ++ndn) {
        std::uint64_t full = 0;
        std::int64_t parents = 0;
        if (ftlm_sector(ctx, nup, ndn, &full, &parents) != FTLM_OK)
          die(std::string("sector build: ") + ftlm_last_error());
        if (full == 0) continue;
        if (full > o.max_sector_dim) {
          std::ostringstream msg;
          msg << "Sector (Nup=" << nup << ", Ndown=" << ndn << ") has full dimension " << full
              << ", which exceeds --max-sector-dim=" << o.max_sector_dim;
          die(msg.str());
        }
        std::vector<int> mxs, mys;
        std::int64_t tot = 0;
        std::size_t active = 0;
        for (int my = 0; my < o.ly; ++my)
          for (int mx = 0; mx < o.lx; ++mx) {
            std::int64_t d = 0;
            if (ftlm_block(ctx, mx, my, &d) != FTLM_OK) die(std::string("block build: ") + ftlm_last_error());
            active += (std::size_t)d;
            if (d == 0) continue;
            mxs.push_back(mx);
            mys.push_back(my);
            tot += d;
          }
        std::vector<double> eig((std::size_t)tot);
        std::vector<std::int64_t> off(mxs.size() + 1, 0);
        if (!mxs.empty() && ftlm_exact_blocks(ctx, (int)mxs.size(), mxs.data(), mys.data(), eig.data(),
                                              off.data()) != FTLM_OK)
          die(std::string("exact diagonalization: ") + ftlm_last_error());
        for (std::size_t b = 0; b < mxs.size(); ++b)
          accumulate(eig.data() + off[b], nullptr, (std::size_t)(off[b + 1] - off[b]), 1.0, nup + ndn, o.beta, mu, t,
                     threads);
        std::cout << "  sector Nup=" << nup << " Ndown=" << ndn << " full_dim=" << full << " parents=" << parents
                  << " active_k_dim_sum=" << active << "\n";
      }
    }
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
