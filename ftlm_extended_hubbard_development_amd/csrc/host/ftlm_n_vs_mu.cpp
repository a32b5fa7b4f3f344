// ftlm_n_vs_mu -- drop-in replacement of the reference executable of the
// same name (src/main.cpp:1459-1839) running the hot path on the GPU engine
// through the C ABI of include/ftlm.h.
//
// Same flags, stdout lines, CSV columns, checkpoint/.meta files and block
// order (N_up, N_down, my, mx) as the reference. Exact blocks of every
// sector and twist are queued and diagonalised in one batched launch. FTLM
// blocks up to --pack-max-dim rows are added to a pack that runs as one
// persistent launch once full (or at the end); larger blocks run on their
// own. Queue and pack also run at the end of every sector when
// checkpointing. Blocks are finished (trace/debug output, checkpoint record,
// accumulation) strictly in the reference's order from an in-order queue, so
// every output keeps the reference's order whatever ran when. With tree
// reductions the start vectors are drawn on the GPU and the thermodynamic
// sums of all blocks run on the GPU at the end. --twist-grid N /
// --twist-list visit every block for all twists, packing the twists of a
// block together (their start vectors are shared: block seeds do not depend
// on the twist).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <deque>
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

double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// One step of the ordered output: a block to finish or a sector line.
struct Item {
  bool is_line = false;
  std::string line;
  BlockKey key{};
  int twist = 0;
  std::int64_t dim = 0;
  int particles = 0;
  bool use_exact = false;
  bool from_ckpt = false;
  bool dbg_here = false;
  Spectrum spec;
  std::vector<double> exact;  // exact eigenvalues (exact, trace or debug blocks)
  std::int64_t pack_idx = -1;
  std::int64_t exact_idx = -1;
  bool need_pack = false;   // waits for its pack's run
  bool need_exact = false;  // waits for the exact queue's run
  bool ready = false;
};

// OUTPUT with a twist tag before the extension: .phixA_phiyB for grid twists,
// .twistI (position in the list) for --twist-list, whose values may not print
// distinctly.
std::string twist_output(const std::string& out, double px, double py, int list_index) {
  std::ostringstream tag;
  if (list_index >= 0) tag << ".twist" << list_index;
  else tag << ".phix" << px << "_phiy" << py;
  const std::size_t slash = out.find_last_of('/');
  const std::size_t dot = out.find_last_of('.');
  if (dot == std::string::npos || (slash != std::string::npos && dot < slash)) return out + tag.str() + ".csv";
  return out.substr(0, dot) + tag.str() + out.substr(dot);
}

// Grand-canonical sums of all stored spectra for every beta in one device
// launch (ftlm_thermo): one Thermo per beta.
std::vector<Thermo> device_sums(ftlm_ctx* c, const std::vector<Spectrum>& specs, const std::vector<double>& betas,
                                const std::vector<double>& mu) {
  std::vector<std::int64_t> off(specs.size() + 1, 0);
  for (std::size_t b = 0; b < specs.size(); ++b) off[b + 1] = off[b] + (std::int64_t)specs[b].eig.size();
  std::vector<double> eig((std::size_t)off.back()), w((std::size_t)off.back(), 1.0), pref(specs.size());
  std::vector<std::int32_t> part(specs.size());
  for (std::size_t b = 0; b < specs.size(); ++b) {
    const Spectrum& s = specs[b];
    std::copy(s.eig.begin(), s.eig.end(), eig.begin() + off[b]);
    if (!s.w.empty()) std::copy(s.w.begin(), s.w.end(), w.begin() + off[b]);
    pref[b] = s.prefactor;
    part[b] = s.particles;
  }
  const std::size_t nmu = mu.size();
  std::vector<double> raw(betas.size() * nmu * 4);
  check(ftlm_thermo(c, (std::int64_t)specs.size(), off.data(), eig.data(), w.data(), part.data(), pref.data(),
                    (int)betas.size(), betas.data(), (int)nmu, mu.data(), raw.data()),
        "thermodynamic sums");
  std::vector<Thermo> all;
  for (std::size_t i = 0; i < betas.size(); ++i) {
    Thermo th(nmu);
    for (std::size_t k = 0; k < nmu; ++k) {
      const double* r = &raw[(i * nmu + k) * 4];
      th.emin[k] = r[0];
      th.z[k] = r[1];
      th.n1[k] = r[2];
      th.n2[k] = r[3];
    }
    all.push_back(th);
  }
  return all;
}

// ... finalised as main.cpp:1770-1788.
std::vector<Results> device_results(ftlm_ctx* c, const std::vector<Spectrum>& specs, const std::vector<double>& betas,
                                    const std::vector<double>& mu, int sites) {
  const std::vector<Thermo> th = device_sums(c, specs, betas, mu);
  std::vector<Results> all;
  for (std::size_t i = 0; i < betas.size(); ++i) all.push_back(finalize(th[i], betas[i], sites));
  return all;
}

// Relative cost of a sector's FTLM blocks for the --stop-after prediction:
// per block R * m * dim * (hop targets per row + 6 vector streams), the
// algorithmic bytes of its Lanczos steps up to a constant factor.
double sector_cost(const Options& o, int nup, int ndn, const std::vector<std::int64_t>& dims) {
  const int L = o.lx * o.ly;
  const double d = 2.0 * (o.tx != 0.0) + 2.0 * (o.ty != 0.0) + 4.0 * (o.tp != 0.0);
  const double h = L > 1 ? d * ((double)nup * (L - nup) + (double)ndn * (L - ndn)) / (L - 1) : 0.0;
  double c = 0.0;
  for (std::int64_t dim : dims)
    if ((std::size_t)dim > o.threshold)
      c += (double)o.samples * (double)std::min<std::int64_t>(o.steps, dim) * (double)dim * (h + 6.0);
  return c;
}

}  // namespace

int main(int argc, char** argv) {
  const auto t_start = std::chrono::steady_clock::now();
  try {
    Options o;
    if (!parse_options(argc, argv, o, false)) return 0;
    const int sites = o.lx * o.ly;
    const bool multi = o.betas.size() > 1;
    // tree mode: keep every spectrum and sum them on the device at the end;
    // serial mode: the reference's streaming host accumulation (bit-identical)
    const bool device_thermo = o.reduction == FTLM_REDUCE_TREE;
    const bool partial = !o.partial_sums.empty();
    // stored spectra: multi-beta, device sums, and resumable partial sums
    // (folded into the running sums at the end of every sector)
    const bool store = multi || device_thermo || partial;
    const std::vector<double> mu = linspace(o.mu_min, o.mu_max, o.mu_count);
    const bool trace = !o.trace_csv.empty();
    const bool dbg = o.dbg_nup >= 0 && o.dbg_ndn >= 0 && o.dbg_mx >= 0 && o.dbg_my >= 0;
    if (!o.checkpoint.empty() && (trace || dbg))
      die("--checkpoint is not currently compatible with trace/debug block output.");
    const bool scan = o.twist_grid > 0 || !o.twist_list.empty();
    if (scan && (!o.checkpoint.empty() || trace || dbg))
      die("--twist-grid/--twist-list are not compatible with --checkpoint or trace/debug block output.");
    if (partial && (trace || dbg)) die("--partial-sums is not compatible with trace/debug block output.");
    std::vector<std::pair<double, double>> twists;
    if (o.twist_grid > 0) {
      for (int iy = 0; iy < o.twist_grid; ++iy)
        for (int ix = 0; ix < o.twist_grid; ++ix)
          twists.push_back({(double)ix / o.twist_grid, (double)iy / o.twist_grid});
    } else if (scan) {
      twists = o.twist_list;
    } else {
      twists.push_back({o.phix, o.phiy});
    }
    const int ntw = (int)twists.size();
    const int nbeta = (int)o.betas.size();
    PartialState ps;
    bool resumed = false;
    if (partial) {
      const std::string want = partial_meta(o, twists);
      resumed = read_partial(o.partial_sums, ps);
      if (resumed) {
        if (ps.meta != want) die("Partial-sum file does not match current run: " + o.partial_sums);
        if (ps.ntwist != ntw || ps.nbeta != nbeta || ps.nmu != (int)mu.size())
          die("Partial-sum file does not match current run: " + o.partial_sums);
      } else {
        ps.meta = want;
        ps.ntwist = ntw;
        ps.nbeta = nbeta;
        ps.nmu = (int)mu.size();
        ps.th.assign((std::size_t)ntw * nbeta, Thermo(mu.size()));
      }
    }
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
    if (o.twist_grid > 0)
      std::cout << "twist grid " << o.twist_grid << "x" << o.twist_grid << " (" << ntw << " twists)\n";
    else if (scan)
      std::cout << "twist list (" << ntw << " twists)\n";
    if (!o.checkpoint.empty())
      std::cout << "checkpoint=" << o.checkpoint << " loaded_blocks=" << ckpt.size() << "\n";
    if (partial)
      std::cout << "partial_sums=" << o.partial_sums << " summed_blocks=" << ps.blocks << " resume_at=Nup="
                << (resumed ? ps.next_nup : 0) << ",Ndown=" << (resumed ? ps.next_ndn : 0) << "\n";

    Ctx ctx;
    const ftlm_model model{o.lx, o.ly, o.tx, o.ty, o.tp, twists[0].first, twists[0].second, o.u, o.v};
    check(ftlm_create(&model, o.device, &ctx.c), "GPU engine unavailable");
    // serial reductions reproduce the reference bit for bit: host-drawn starts
    check(ftlm_set_start_source(ctx.c, device_thermo && !o.host_starts ? FTLM_START_DEVICE : FTLM_START_HOST),
          "start vectors");
    check(ftlm_set_vector_placement(ctx.c, o.vector_placement, (std::uint64_t)256 << 20), "vector placement");
    int cur_tw = 0;
    auto use_twist = [&](int t) {
      if (t == cur_tw) return;
      check(ftlm_set_twist(ctx.c, twists[t].first, twists[t].second), "twist");
      cur_tw = t;
    };

    std::vector<std::vector<Spectrum>> stored(ntw);
    std::vector<Thermo> single(ntw, Thermo(mu.size()));
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
    auto tj_rec = [&](const std::string& body) {
      if (!tj) return;
      tj << (first_tj ? "" : ",\n") << "{" << body << "}" << std::flush;
      first_tj = false;
    };
    const int L = sites;
    const std::size_t R = (std::size_t)o.samples;

    // ---- finishing a block (the tail of the reference's block loop) ----
    auto finish_item = [&](Item& it) {
      if (it.is_line) {
        std::cout << it.line;
        return;
      }
      const int nup = it.key.nup, ndn = it.key.ndn, mx = it.key.mx, my = it.key.my;
      const int particles = it.particles;
      const std::int64_t dim = it.dim;
      Spectrum& spec = it.spec;
      const int t = it.twist;
      if (it.from_ckpt) {
        if (store) stored[t].push_back(std::move(spec));
        else
          accumulate(spec.eig.data(), spec.w.empty() ? nullptr : spec.w.data(), spec.eig.size(), spec.prefactor,
                     spec.particles, o.beta, mu, single[t], threads);
        return;
      }
      const bool use_exact = it.use_exact;
      const double* exact_eig = it.exact.empty() ? nullptr : it.exact.data();
      const std::size_t exact_n = it.exact.size();
      if (use_exact) {
        spec.eig = it.exact;
        spec.prefactor = 1.0;
      }
      const double sumw = use_exact ? 1.0 : std::accumulate(spec.w.begin(), spec.w.end(), 0.0);
      if (trace) {
        for (double m : mu) {
          const BlockThermo used = use_exact ? raw_thermo(exact_eig, nullptr, exact_n, 1.0, particles, o.beta, m)
                                             : raw_thermo(spec.eig.data(), spec.w.data(), spec.eig.size(),
                                                          (double)dim, particles, o.beta, m);
          const BlockThermo ex = raw_thermo(exact_eig, nullptr, exact_n, 1.0, particles, o.beta, m);
          const double ratio = ex.z > 0.0 ? used.z / ex.z : 0.0;
          tout << nup << "," << ndn << "," << mx << "," << my << "," << particles << "," << dim << "," << m << ","
               << (use_exact ? "exact" : "ftlm") << "," << used.z << "," << ex.z << "," << ratio << ","
               << used.n_total << "," << ex.n_total << "," << (used.n_total - ex.n_total) << "," << used.n2_total
               << "," << ex.n2_total << "," << (used.n2_total - ex.n2_total) << "," << used.min_shifted << ","
               << ex.min_shifted << "," << sumw << "\n";
        }
      }
      if (it.dbg_here) {
        std::cout << "Selected block debug Nup=" << nup << " Ndown=" << ndn << " mx=" << mx << " my=" << my
                  << " basis_dim=" << dim << " method=" << (use_exact ? "exact" : "ftlm")
                  << " sum_sample_weights=" << sumw << "\n";
        for (double m : mu) {
          const BlockThermo used = use_exact ? raw_thermo(exact_eig, nullptr, exact_n, 1.0, particles, o.beta, m)
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
                 << ex.z << "," << ratio << "," << f.n_total << "," << ex.n_total << "," << (f.n_total - ex.n_total)
                 << "," << f.n2_total << "," << ex.n2_total << "," << (f.n2_total - ex.n2_total) << ","
                 << f.min_shifted << "," << ex.min_shifted << "," << sw << "\n";
          }
          wrote_dbg = true;
        }
      }
      if (multi || !o.checkpoint.empty()) append_record(o.checkpoint, it.key, spec);
      if (store)
        stored[t].push_back(std::move(spec));
      else
        accumulate(spec.eig.data(), spec.w.empty() ? nullptr : spec.w.data(), spec.eig.size(), spec.prefactor,
                   particles, o.beta, mu, single[t], threads);
    };

    std::deque<Item> queue;

    // ---- the pack of small FTLM blocks ----
    const bool packing = o.pack_max_dim > 0 && !o.per_vector;
    const int pack_method = o.reduction == FTLM_REDUCE_SERIAL ? 0 : 1;  // Jacobi (bit-exact) or QL Ritz
    const std::uint64_t budget = (std::uint64_t)(o.pack_budget_gb * 1073741824.0);
    std::vector<std::size_t> pending;  // push positions of the blocks in the current pack
    std::int64_t pack_entries = 0;
    bool tuned = o.pack_width > 0 && o.pack_wg > 0;
    if (packing) check(ftlm_pack_begin(ctx.c, o.samples, o.steps, budget), "pack");
    std::size_t popped = 0;  // items removed from the queue front so far
    std::size_t pushed = 0;
    auto push = [&](Item&& it) {
      queue.push_back(std::move(it));
      return pushed++;
    };
    auto drain = [&]() {
      while (!queue.empty() && queue.front().ready) {
        finish_item(queue.front());
        queue.pop_front();
        ++popped;
      }
    };
    auto flush_pack = [&]() {
      if (!packing || pack_entries == 0) return;
      const auto t0 = std::chrono::steady_clock::now();
      if (!tuned) {
        int bw = 0, bg = 0, n = 0;
        double bms = 0.0;
        // two Lanczos steps per candidate (the first one also loads the start
        // vectors): enough to rank (width, grid), a small share of the run
        check(ftlm_pack_tune(ctx.c, 2, &bw, &bg, &bms, &n, nullptr, nullptr, nullptr), "pack tuning");
        tuned = true;
        std::ostringstream rec;
        rec << "\"method\":\"pack_tune\",\"width\":" << bw << ",\"wg_per_cu\":" << bg << ",\"ms\":" << ms_since(t0);
        tj_rec(rec.str());
      }
      check(ftlm_pack_run(ctx.c, o.pack_width, o.pack_wg, pack_method), "packed FTLM");
      for (std::size_t pos : pending) {
        Item& it = queue[pos - popped];
        const std::size_t m = std::min<std::size_t>((std::size_t)o.steps, (std::size_t)it.dim);
        it.spec.eig.resize(R * m);
        it.spec.w.resize(R * m);
        std::int64_t cnt = 0;
        check(ftlm_pack_result(ctx.c, it.pack_idx, it.spec.eig.data(), it.spec.w.data(), &cnt), "packed FTLM");
        it.spec.eig.resize((std::size_t)cnt);
        it.spec.w.resize((std::size_t)cnt);
        it.need_pack = false;
        it.ready = !it.need_exact;
      }
      ftlm_stats st{};
      ftlm_last_stats(ctx.c, &st);
      std::ostringstream rec;
      rec << "\"method\":\"pack\",\"blocks\":" << pack_entries << ",\"gpu_ms\":" << st.ms_total
          << ",\"ms\":" << ms_since(t0) << ",\"bytes_apply\":" << st.bytes_apply;
      tj_rec(rec.str());
      pending.clear();
      pack_entries = 0;
      check(ftlm_pack_begin(ctx.c, o.samples, o.steps, budget), "pack");
      drain();
    };

    // ---- the queue of exact blocks (one batched eigensolve) ----
    std::vector<std::size_t> ex_pending;  // push positions waiting for the queue
    std::int64_t exact_entries = 0;
    check(ftlm_exact_queue_begin(ctx.c, 0), "exact diagonalization");
    auto flush_exact = [&]() {
      if (exact_entries == 0) return;
      const auto t0 = std::chrono::steady_clock::now();
      check(ftlm_exact_queue_run(ctx.c), "exact diagonalization");
      for (std::size_t pos : ex_pending) {
        Item& it = queue[pos - popped];
        std::int64_t cnt = 0;
        check(ftlm_exact_queue_result(ctx.c, it.exact_idx, nullptr, &cnt), "exact diagonalization");
        it.exact.resize((std::size_t)cnt);
        check(ftlm_exact_queue_result(ctx.c, it.exact_idx, it.exact.data(), &cnt), "exact diagonalization");
        it.need_exact = false;
        it.ready = !it.need_pack;
      }
      tj_rec("\"method\":\"exact_batch\",\"blocks\":" + std::to_string(exact_entries) +
             ",\"ms\":" + std::to_string(ms_since(t0)));
      ex_pending.clear();
      exact_entries = 0;
      check(ftlm_exact_queue_begin(ctx.c, 0), "exact diagonalization");
      drain();
    };

    // ---- resumable runs: fold finished sectors into the running sums ----
    const double base_elapsed = ps.elapsed_s;
    const int base_runs = ps.runs;
    auto seconds = [&]() { return ms_since(t_start) / 1000.0; };
    auto fold = [&]() {
      for (int t = 0; t < ntw; ++t) {
        if (stored[t].empty()) continue;
        if (device_thermo) {
          const std::vector<Thermo> part = device_sums(ctx.c, stored[t], o.betas, mu);
          for (int b = 0; b < nbeta; ++b) merge_thermo(ps.th[(std::size_t)t * nbeta + b], part[b], o.betas[b]);
        } else {
          // the reference's streaming sums, continued block by block
          for (int b = 0; b < nbeta; ++b)
            for (const Spectrum& sp : stored[t])
              accumulate(sp.eig.data(), sp.w.empty() ? nullptr : sp.w.data(), sp.eig.size(), sp.prefactor,
                         sp.particles, o.betas[b], mu, ps.th[(std::size_t)t * nbeta + b], threads);
        }
        ps.blocks += (std::int64_t)stored[t].size();
        stored[t].clear();
      }
    };
    auto save_partial = [&](int next_nup, int next_ndn) {
      ps.next_nup = next_nup;
      ps.next_ndn = next_ndn;
      ps.runs = base_runs + 1;
      ps.elapsed_s = base_elapsed + seconds();
      write_partial(o.partial_sums, ps);
    };
    double work_here = 0.0;      // sector_cost of the sectors run by this process
    double t_first_sector = -1.0;

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
        // block dimensions (twist independent)
        std::vector<std::int64_t> dims(L);
        for (int my = 0; my < o.ly; ++my)
          for (int mx = 0; mx < o.lx; ++mx) check(ftlm_block(ctx.c, mx, my, &dims[my * o.lx + mx]), "block build");
        const bool summed_before =
            partial && resumed && (nup < ps.next_nup || (nup == ps.next_nup && ndn < ps.next_ndn));
        if (!summed_before && o.stop_after > 0.0 && work_here > 0.0) {
          const double el = seconds();
          const double cost = sector_cost(o, nup, ndn, dims) * (double)ntw;
          const double pred = cost * ((el - t_first_sector) / work_here);
          if (el + pred > o.stop_after) {
            if (partial) save_partial(nup, ndn);
            if (tj) tj << "\n]\n";
            std::cout << "Stopped before sector Nup=" << nup << " Ndown=" << ndn << " after " << el
                      << " s (--stop-after " << o.stop_after << ", next sector predicted " << pred
                      << " s); run the same command again to continue\n";
            return 0;
          }
        }
        if (summed_before) {
          std::size_t active = 0;
          for (std::int64_t d : dims) active += (std::size_t)d;
          std::ostringstream line;
          line << "  sector Nup=" << nup << " Ndown=" << ndn << " full_dim=" << full << " parents=" << parents
               << " active_k_dim_sum=" << active << "\n";
          Item li;
          li.is_line = true;
          li.line = line.str();
          li.ready = true;
          push(std::move(li));
          drain();
          continue;
        }
        if (t_first_sector < 0.0) t_first_sector = seconds();
        work_here += sector_cost(o, nup, ndn, dims) * (double)ntw;
        std::size_t active = 0;
        for (int my = 0; my < o.ly; ++my) {
          for (int mx = 0; mx < o.lx; ++mx) {
            const int k = my * o.lx + mx;
            const std::int64_t dim = dims[k];
            active += (std::size_t)dim;
            if (dim == 0) continue;
            for (int t = 0; t < ntw; ++t) {
              Item it;
              it.key = BlockKey{nup, ndn, mx, my};
              it.twist = t;
              it.dim = dim;
              it.particles = nup + ndn;
              const auto hit = ckpt.find(it.key);
              if (hit != ckpt.end()) {
                const Spectrum& s = hit->second;
                if (s.dim != (std::uint64_t)dim || s.particles != it.particles)
                  die("Checkpoint block metadata does not match current run.");
                it.spec = s;
                it.from_ckpt = true;
                it.ready = true;
                push(std::move(it));
                continue;
              }
              it.use_exact = (std::size_t)dim <= o.threshold;
              it.dbg_here = dbg && nup == o.dbg_nup && ndn == o.dbg_ndn && mx == o.dbg_mx && my == o.dbg_my;
              it.spec.particles = it.particles;
              it.spec.dim = (std::uint64_t)dim;
              it.spec.exact = it.use_exact;
              use_twist(t);
              // exact eigenvalues (exact blocks, and trace/debug references)
              // come from the queued batch eigensolve
              it.need_exact = it.use_exact || trace || it.dbg_here;
              if (it.need_exact) {
                int rc = ftlm_exact_queue_add(ctx.c, mx, my, &it.exact_idx);
                if (rc == FTLM_PACK_FULL) {
                  flush_exact();
                  rc = ftlm_exact_queue_add(ctx.c, mx, my, &it.exact_idx);
                }
                check(rc == FTLM_PACK_FULL ? FTLM_ERR_LIMIT : rc, "exact diagonalization");
                ++exact_entries;
              }
              if (it.use_exact) {
                ex_pending.push_back(push(std::move(it)));
                continue;
              }
              it.spec.prefactor = (double)dim;
              check(ftlm_block(ctx.c, mx, my, nullptr), "block build");
              const std::uint64_t bseed = ftlm_block_seed(o.seed, nup, ndn, mx, my);
              if (packing && dim <= o.pack_max_dim) {
                int rc = ftlm_pack_add(ctx.c, bseed, &it.pack_idx);
                if (rc == FTLM_PACK_FULL) {
                  flush_pack();
                  rc = ftlm_pack_add(ctx.c, bseed, &it.pack_idx);
                }
                check(rc == FTLM_PACK_FULL ? FTLM_ERR_LIMIT : rc, "pack");
                ++pack_entries;
                it.need_pack = true;
                const bool ex = it.need_exact;
                const std::size_t pos = push(std::move(it));
                pending.push_back(pos);
                if (ex) ex_pending.push_back(pos);
                continue;
              }
              const auto t0 = std::chrono::steady_clock::now();
              const std::size_t m = std::min<std::size_t>((std::size_t)o.steps, (std::size_t)dim);
              it.spec.eig.resize(R * m);
              it.spec.w.resize(R * m);
              std::int64_t cnt = 0;
              check(ftlm_ftlm_block(ctx.c, o.samples, o.steps, bseed, o.reduction, o.per_vector ? 1 : 0,
                                    it.spec.eig.data(), it.spec.w.data(), &cnt),
                    "FTLM block");
              it.spec.eig.resize((std::size_t)cnt);
              it.spec.w.resize((std::size_t)cnt);
              it.ready = !it.need_exact;
              tj_rec("\"nup\":" + std::to_string(nup) + ",\"ndn\":" + std::to_string(ndn) + ",\"mx\":" +
                     std::to_string(mx) + ",\"my\":" + std::to_string(my) + ",\"dim\":" + std::to_string(dim) +
                     ",\"method\":\"ftlm\",\"ms\":" + std::to_string(ms_since(t0)));
              const bool ex = it.need_exact;
              const std::size_t pos = push(std::move(it));
              if (ex) ex_pending.push_back(pos);
            }
            drain();
          }
        }
        // a checkpointed (or partial-sum) run finishes every sector before
        // the next one starts, so an interrupted run keeps its progress
        if (!o.checkpoint.empty() || partial) {
          flush_pack();
          flush_exact();
        }
        std::ostringstream line;
        line << "  sector Nup=" << nup << " Ndown=" << ndn << " full_dim=" << full << " parents=" << parents
             << " active_k_dim_sum=" << active << "\n";
        Item li;
        li.is_line = true;
        li.line = line.str();
        li.ready = true;
        push(std::move(li));
        drain();
        if (partial) {
          if (!queue.empty()) die("internal error: unfinished blocks at the end of a sector");
          fold();
          save_partial(ndn < L ? nup : nup + 1, ndn < L ? ndn + 1 : 0);
        }
      }
    }
    flush_pack();
    flush_exact();
    drain();
    if (partial) {
      fold();
      save_partial(L + 1, 0);
      std::cout << "partial sums: " << ps.blocks << " blocks summed over " << ps.runs << " run(s), "
                << ps.elapsed_s << " s of wall time in total\n";
    }
    if (!queue.empty()) die("internal error: unfinished blocks");
    if (tj) tj << "\n]\n";
    if (dbg && !o.debug_csv.empty() && !wrote_dbg)
      die("Selected block debug target was not found among active FTLM blocks.");

    for (int t = 0; t < ntw; ++t) {
      const std::string out =
          scan ? twist_output(o.output, twists[t].first, twists[t].second, o.twist_list.empty() ? -1 : t) : o.output;
      if (scan) std::cout << "twist phix=" << twists[t].first << " phiy=" << twists[t].second << "\n";
      std::vector<Results> all;
      if (partial) {
        for (int b = 0; b < nbeta; ++b) all.push_back(finalize(ps.th[(std::size_t)t * nbeta + b], o.betas[b], sites));
      } else if (device_thermo) {
        all = device_results(ctx.c, stored[t], o.betas, mu, sites);
      } else if (multi) {
        for (double beta : o.betas) {
          Thermo th(mu.size());
          for (const Spectrum& s : stored[t])
            accumulate(s.eig.data(), s.w.empty() ? nullptr : s.w.data(), s.eig.size(), s.prefactor, s.particles,
                       beta, mu, th, threads);
          all.push_back(finalize(th, beta, sites));
        }
      } else {
        all.push_back(finalize(single[t], o.beta, sites));
      }
      for (std::size_t b = 0; b < all.size(); ++b) {
        const Results& r = all[b];
        for (std::size_t i = 0; i < mu.size(); ++i)
          std::cout << "beta=" << std::setw(12) << o.betas[b] << "  mu=" << std::setw(12) << mu[i]
                    << "  n=" << std::setw(12) << r.n[i] << "  kappa=" << std::setw(12) << r.kappa[i] << "\n";
      }
      if (multi) write_multi_beta(out, o.betas, mu, all);
      else write_results(out, mu, all[0]);
      std::cout << "Wrote " << out << "\n";
    }
    return 0;
  } catch (const std::exception& ex) {
    std::cerr << "Error: " << ex.what() << "\n";
    return 1;
  }
}
