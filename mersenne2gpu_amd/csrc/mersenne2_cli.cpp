// mersenne2 <p> [ll|prp] [report_every] [-d id] [-R sec] [-W wg] [-CW wg] [-B carry_pairs]
//           [-LS cap] [-N max_iters] [--show-defaults] [--enable-chunk64] [--debug-plan]
//           [--debug-stage-selftest] [--debug-cppint-every N] [--debug-cppint-max-iter N]
//           [--res64-at k1,k2,...] [--dump-residue path] [--checkpoint path] [--resume path]
//           [--gerbicz block] [--profile-kernels] [--profile-every N]
//
// Drop-in for the reference's command line (mersenne2_gpu.cpp:3618-4960): same
// grammar, defaults, stdout lines and exit codes, on top of the libm2 C ABI. It
// adds a final "Res64: <hex>" line (the reference prints no residue), the
// --res64-at / --dump-residue options, checkpoints, and a Gerbicz error check
// for PRP runs (--gerbicz: blocks of that many iterations, a check every ten).
// --profile-kernels / --profile-every N print per-kernel device-time totals
// the way the reference's power-only prototype does
// (gf61_ntt_prp_opencl_inplace_nobitrev_fused_powonly_blockcarry.cpp:645-698,
// :1277-1313): a summary every N iterations (default: the report interval)
// and a final one.
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <iomanip>
#include <iostream>
#include <limits>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "mersenne2.h"

namespace {

const char* USAGE_TAIL =
    " <p> [ll|prp] [report_every] [-d device_id] [-R report_seconds] [-W fft_wg] [-CW carry_wg] [-B carry_pairs] "
    "[-LS local_stage_cap] [-N max_iters] [--show-defaults] [--enable-chunk64] [--debug-plan] "
    "[--debug-stage-selftest] [--debug-cppint-every N] [--debug-cppint-max-iter N] [--res64-at k1,k2,...] "
    "[--dump-residue path] [--checkpoint path] [--resume path] [--gerbicz block] [--profile-kernels] "
    "[--profile-every N]\n";

int die(const char* what, int code) {
  std::cerr << what << " failed with error " << code << ": " << m2_last_error() << std::endl;
  return 1;
}

// Per-kernel totals in order of first launch.
struct Profile {
  std::vector<std::string> order;
  std::map<std::string, std::pair<double, uint64_t>> acc;  // name -> (total ms, launches)
};

int profile_run(m2_test* t, uint64_t iters, Profile& pr) {
  constexpr int MAXK = 16;
  double ms[MAXK];
  uint64_t nl[MAXK];
  int cnt = 0;
  char names[1024] = {0};
  int rc = m2_test_kernel_profile(t, iters, ms, nl, MAXK, &cnt, names, sizeof names);
  if (rc != M2_OK) return rc;
  std::stringstream ss(names);
  std::string nm;
  for (int k = 0; k < cnt && std::getline(ss, nm); ++k) {
    if (!pr.acc.count(nm)) pr.order.push_back(nm);
    pr.acc[nm].first += ms[k];
    pr.acc[nm].second += nl[k];
  }
  return M2_OK;
}

void profile_print(const Profile& pr, const std::string& title) {
  if (pr.order.empty()) {
    std::cout << title << ": no event timing data collected.\n";
    return;
  }
  double total = 0;
  for (const auto& nm : pr.order) total += pr.acc.at(nm).first;
  if (total <= 0) {
    std::cout << title << ": events were recorded but their timings are unavailable.\n";
    return;
  }
  std::vector<std::pair<std::string, std::pair<double, uint64_t>>> rows;
  for (const auto& nm : pr.order) rows.push_back({nm, pr.acc.at(nm)});
  std::stable_sort(rows.begin(), rows.end(), [](const auto& a, const auto& b) { return a.second.first > b.second.first; });
  std::cout << title << ":\n";
  for (const auto& r : rows)
    std::cout << "  " << r.first << ": " << std::fixed << std::setprecision(3) << r.second.first << " ms total ("
              << std::setprecision(1) << 100.0 * r.second.first / total << "%, launches=" << r.second.second << ")\n";
  std::cout << std::flush;
}

std::string hex16(uint64_t v) {
  char b[32];
  snprintf(b, sizeof b, "%016llX", (unsigned long long)v);
  return b;
}

}  // namespace

int main(int argc, char* argv[]) {
  if (argc < 2) {
    std::cerr << "Usage: " << argv[0] << USAGE_TAIL;
    return 1;
  }
  const uint32_t p = static_cast<uint32_t>(std::strtoul(argv[1], nullptr, 10));
  int mode = M2_MODE_LL;  // the reference defaults to LL (mersenne2_gpu.cpp:3625)
  uint32_t report_every = 1000u;
  bool report_every_set = false;
  double report_seconds = 2.0;
  uint32_t device_index = 0, wg_override = 0, carry_wg_override = 0, carry_pairs_override = 0;
  uint32_t local_stage_cap_override = 0, max_iters_override = 0;
  bool show_defaults = false, debug_plan = false, debug_stage_selftest = false, enable_chunk64 = false;
  uint32_t debug_cppint_every = 0, debug_cppint_max_iter = 0;
  std::vector<uint64_t> res64_at;
  std::string dump_path, checkpoint_path, resume_path;
  uint32_t gerbicz_block = 0;
  bool profile_kernels = false;
  uint32_t profile_every = 0;

  auto need = [&](int i, const char* msg) -> bool {
    if (i + 1 >= argc) {
      std::cerr << msg << "\n";
      return false;
    }
    return true;
  };
  for (int i = 2; i < argc; ++i) {
    const std::string arg = argv[i];
    if (arg == "ll" || arg == "prp") {
      mode = (arg == "ll") ? M2_MODE_LL : M2_MODE_PRP;
    } else if (arg == "-d") {
      if (!need(i, "Missing device id after -d")) return 1;
      device_index = (uint32_t)std::strtoul(argv[++i], nullptr, 10);
    } else if (arg == "-R") {
      if (!need(i, "Missing seconds after -R")) return 1;
      report_seconds = std::max(0.0, std::atof(argv[++i]));
    } else if (arg == "-W") {
      if (!need(i, "Missing work-group size after -W")) return 1;
      wg_override = (uint32_t)std::strtoul(argv[++i], nullptr, 10);
    } else if (arg == "-CW") {
      if (!need(i, "Missing carry work-group size after -CW")) return 1;
      carry_wg_override = (uint32_t)std::strtoul(argv[++i], nullptr, 10);
    } else if (arg == "-B") {
      if (!need(i, "Missing carry_pairs after -B")) return 1;
      carry_pairs_override = (uint32_t)std::strtoul(argv[++i], nullptr, 10);
    } else if (arg == "-LS") {
      if (!need(i, "Missing local stage cap after -LS")) return 1;
      local_stage_cap_override = (uint32_t)std::strtoul(argv[++i], nullptr, 10);
    } else if (arg == "-N") {
      if (!need(i, "Missing max iterations after -N")) return 1;
      max_iters_override = (uint32_t)std::strtoul(argv[++i], nullptr, 10);
    } else if (arg == "--show-defaults") {
      show_defaults = true;
    } else if (arg == "--debug-plan") {
      debug_plan = true;
    } else if (arg == "--debug-stage-selftest") {
      debug_stage_selftest = true;
    } else if (arg == "--enable-chunk64") {
      enable_chunk64 = true;
    } else if (arg == "--debug-cppint-every") {
      if (!need(i, "Missing value after --debug-cppint-every")) return 1;
      debug_cppint_every = (uint32_t)std::strtoul(argv[++i], nullptr, 10);
    } else if (arg == "--debug-cppint-max-iter") {
      if (!need(i, "Missing value after --debug-cppint-max-iter")) return 1;
      debug_cppint_max_iter = (uint32_t)std::strtoul(argv[++i], nullptr, 10);
    } else if (arg == "--res64-at") {
      if (!need(i, "Missing iteration list after --res64-at")) return 1;
      std::stringstream ss(argv[++i]);
      std::string tok;
      while (std::getline(ss, tok, ',')) {
        if (!tok.empty()) res64_at.push_back(std::strtoull(tok.c_str(), nullptr, 10));
      }
      std::sort(res64_at.begin(), res64_at.end());
    } else if (arg == "--dump-residue") {
      if (!need(i, "Missing path after --dump-residue")) return 1;
      dump_path = argv[++i];
    } else if (arg == "--checkpoint") {
      if (!need(i, "Missing path after --checkpoint")) return 1;
      checkpoint_path = argv[++i];
    } else if (arg == "--resume") {
      if (!need(i, "Missing path after --resume")) return 1;
      resume_path = argv[++i];
    } else if (arg == "--gerbicz") {
      if (!need(i, "Missing block length after --gerbicz")) return 1;
      gerbicz_block = (uint32_t)std::strtoul(argv[++i], nullptr, 10);
      if (gerbicz_block == 0) {
        std::cerr << "--gerbicz needs a block length > 0\n";
        return 1;
      }
    } else if (arg == "--profile-kernels") {
      profile_kernels = true;
    } else if (arg == "--profile-every") {
      if (!need(i, "Missing value after --profile-every")) return 1;
      profile_every = (uint32_t)std::strtoul(argv[++i], nullptr, 10);
    } else if (!report_every_set) {
      report_every = (uint32_t)std::strtoul(arg.c_str(), nullptr, 10);
      report_every_set = true;
    } else {
      std::cerr << "Unknown argument: " << arg << "\n";
      std::cerr << "Usage: " << argv[0] << USAGE_TAIL;
      return 1;
    }
  }
  (void)report_seconds;  // accepted, as in the reference, where it has no effect
  (void)enable_chunk64;  // accepted; this engine has no opt-in chunk64 kernels

  if (debug_cppint_every != 0u || debug_cppint_max_iter != 0u) {
    std::cout << "debug_cppint_backend=disabled (no exact bigint backend in this engine)\n";
    std::cout << "debug_cppint disabled at runtime; use --res64-at to compare residues with an exact reference\n";
    debug_cppint_every = debug_cppint_max_iter = 0;
  }

  int ln = 0;
  if (m2_transform_log2(p, &ln) != M2_OK) {
    std::cerr << "Unsupported exponent " << p << ": " << m2_last_error() << "\n";
    return 1;
  }
  const uint64_t n = uint64_t(1) << ln, h = n / 2;
  std::cout << "p=" << p << ", ln=" << ln << ", transform=" << h << " (n=" << n << ")\n";

  m2_ctx* ctx = nullptr;
  int rc = m2_open((int)device_index, &ctx);
  if (rc != M2_OK) {
    std::cerr << m2_last_error() << "\n";
    return 1;
  }
  char name[256] = {0};
  m2_device_name(ctx, name, sizeof name);
  std::cout << "device=" << device_index << " (" << name << ")\n";
  uint32_t cu = 0, max_wg = 0, lds = 0;
  uint64_t mem = 0;
  m2_device_props(ctx, &cu, &mem, &max_wg, &lds);

  // Defaults of the AMD branch (mersenne2_gpu.cpp:3779-3788): report every 1000
  // iterations, 2000 from h = 2^20 on.
  if (!report_every_set) report_every = (h >= (uint64_t(1) << 20)) ? 2000u : 1000u;
  const uint32_t auto_wg = 256, auto_carry_pairs = 4;
  const std::string profile = "mi355x-gfx950";
  uint32_t wg = auto_wg, carry_wg = auto_wg;
  if (wg_override) wg = carry_wg = std::max<uint32_t>(1, std::min(wg_override, max_wg));
  if (carry_wg_override) carry_wg = std::max<uint32_t>(1, std::min(carry_wg_override, max_wg));
  const uint32_t local_stage_cap = local_stage_cap_override ? std::min(max_wg, local_stage_cap_override) : 2048u;
  if (show_defaults) {
    std::cout << "vendor=AMD, compute_units=" << cu << ", max_wg=" << max_wg << ", local_mem=" << lds << "\n"
              << "auto_wg=" << auto_wg << ", auto_carry_wg=" << auto_wg << ", auto_carry_pairs=" << auto_carry_pairs
              << ", auto_profile=" << profile << "\n"
              << "effective_wg=" << wg << ", effective_carry_wg=" << carry_wg
              << ", effective_local_stage_cap=" << local_stage_cap << ", effective_max_iters=" << max_iters_override
              << "\n";
    m2_close(ctx);
    return 0;
  }
  if (debug_stage_selftest)
    std::cout << "stage selftest: skipped (this engine has no host-emulated fused stages)\n";

  m2_opts opts{};
  opts.wg = wg;
  opts.carry_wg = carry_wg;
  opts.carry_pairs = carry_pairs_override;
  opts.local_stage_cap = local_stage_cap_override;
  m2_test* t = nullptr;
  rc = m2_test_create(ctx, p, mode, &opts, &t);
  if (rc != M2_OK) return die("m2_test_create", rc);

  char plan[4096] = {0};
  m2_test_plan(t, plan, sizeof plan);
  // carry geometry: segments of the outer pass (or of the single workgroup)
  uint64_t nseg = h / 4, seg = 4;
  {
    const char* s = strstr(plan, "carry segments=");
    if (s) {
      nseg = std::strtoull(s + 15, nullptr, 10);
      seg = nseg ? h / nseg : 4;
    }
  }
  std::cout << "carry_blocks=" << nseg << ", avg_pairs_per_block=" << std::fixed << std::setprecision(2)
            << (double)h / (double)nseg << ", carry_target_pairs=" << seg << ", tuned_wg=" << wg
            << ", carry_wg=" << carry_wg << ", local_stage_cap=" << local_stage_cap << ", auto_profile=" << profile;
  if (wg_override != 0u || carry_pairs_override != 0u) std::cout << " (manual override)";
  std::cout << "\n";
  if (debug_plan) std::cout << plan;

  uint64_t total = 0;
  m2_test_total_iters(t, &total);
  if (max_iters_override != 0u) total = std::min<uint64_t>(total, max_iters_override);
  if (max_iters_override != 0u) std::cout << "benchmark_iters=" << total << "\n";

  uint64_t start = 0;
  if (!resume_path.empty()) {
    rc = m2_checkpoint_load(t, resume_path.c_str());
    if (rc != M2_OK) return die("m2_checkpoint_load", rc);
    m2_test_iteration(t, &start);
    if (start > total) start = total;
    std::cout << "resumed from " << resume_path << " at iteration " << start << "\n";
  }
  const auto t0 = std::chrono::steady_clock::now();
  auto report = [&](uint64_t iter) {
    const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    const double pct = total ? 100.0 * (double)iter / (double)total : 100.0;
    const double itps = el > 0 ? (double)(iter - start) / el : 0.0;
    const double eta = itps > 0 ? (double)(total - iter) / itps : std::numeric_limits<double>::infinity();
    std::cout << "iter " << iter << "/" << total << " (" << std::fixed << std::setprecision(1) << pct << "%)"
              << ", elapsed " << std::setprecision(2) << el << " s, it/s " << std::setprecision(1) << itps;
    if (std::isfinite(eta)) std::cout << ", ETA " << std::setprecision(1) << eta << " s";
    std::cout << "\n" << std::flush;
  };
  report(start);
  const uint64_t batch = std::max<uint32_t>(1u, report_every ? report_every : 1000u);
  // the Gerbicz check runs ten blocks per call (one check each) and reports per call
  const bool gerbicz = gerbicz_block && mode == M2_MODE_PRP;
  constexpr uint32_t GERBICZ_BLOCKS = 10;
  uint64_t gerbicz_errors = 0;
  // per-kernel profile: kernels timed by events (the Gerbicz path runs untimed)
  if (profile_kernels && gerbicz) std::cout << "profile: --gerbicz runs are not profiled\n";
  const uint64_t prof_every = (profile_kernels && !gerbicz) ? (profile_every ? profile_every : batch) : 0;
  Profile prof;
  uint64_t iter = start;
  size_t next_cp = 0;
  while (iter < total) {
    uint64_t chunk_end = gerbicz ? std::min<uint64_t>(total, iter + (uint64_t)gerbicz_block * GERBICZ_BLOCKS)
                                 : std::min<uint64_t>(total, (iter / batch + 1) * batch);
    if (prof_every) chunk_end = std::min<uint64_t>(chunk_end, (iter / prof_every + 1) * prof_every);
    if (next_cp < res64_at.size() && res64_at[next_cp] > iter && res64_at[next_cp] < chunk_end)
      chunk_end = res64_at[next_cp];
    if (gerbicz) {
      uint64_t e = 0;
      rc = m2_test_run_checked(t, chunk_end - iter, gerbicz_block, GERBICZ_BLOCKS, &e);
      gerbicz_errors += e;
      if (rc != M2_OK) return die("m2_test_run_checked", rc);
    } else if (prof_every) {
      rc = profile_run(t, chunk_end - iter, prof);
      if (rc != M2_OK) return die("m2_test_kernel_profile", rc);
    } else {
      rc = m2_test_run(t, chunk_end - iter);
      if (rc == M2_OK) rc = m2_test_sync(t);
      if (rc != M2_OK) return die("m2_test_run", rc);
    }
    iter = chunk_end;
    if (prof_every && (iter % prof_every == 0 || iter == total))
      profile_print(prof, "Kernel profile summary at iter " + std::to_string(iter));
    while (next_cp < res64_at.size() && res64_at[next_cp] <= iter) {
      if (res64_at[next_cp] == iter) {
        uint64_t r = 0;
        if (m2_test_res64(t, &r) != M2_OK) return die("m2_test_res64", -1);
        std::cout << "Res64@" << iter << ": " << hex16(r) << "\n";
      }
      ++next_cp;
    }
    if (iter % batch == 0 || iter == total || gerbicz) {
      report(iter);
      if (!checkpoint_path.empty()) {
        rc = m2_checkpoint_save(t, checkpoint_path.c_str());
        if (rc != M2_OK) return die("m2_checkpoint_save", rc);
      }
    }
  }
  rc = m2_test_sync(t);
  if (rc != M2_OK) return die("m2_test_sync", rc);
  int prime = 0;
  if (m2_test_is_prime(t, &prime) != M2_OK) return die("m2_test_is_prime", -1);
  uint64_t r64 = 0;
  if (m2_test_res64(t, &r64) != M2_OK) return die("m2_test_res64", -1);
  const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  const double itps = el > 0 ? (double)(total - start) / el : 0.0;
  if (gerbicz) std::cout << "gerbicz: block " << gerbicz_block << ", " << gerbicz_errors << " failed checks\n";
  std::cout << "2^" << p << " - 1 is "
            << (prime ? (mode == M2_MODE_LL ? "prime" : "a probable prime") : "composite") << ", elapsed "
            << std::fixed << std::setprecision(2) << el << " s, it/s " << std::setprecision(1) << itps << "\n";
  if (prof_every) profile_print(prof, "Kernel profile summary (final)");
  std::cout << "Res64: " << hex16(r64) << "\n";
  if (!dump_path.empty()) {
    std::vector<uint8_t> buf((p + 7) / 8);
    if (m2_test_residue(t, buf.data(), buf.size()) != M2_OK) return die("m2_test_residue", -1);
    FILE* f = fopen(dump_path.c_str(), "wb");
    if (!f || fwrite(buf.data(), 1, buf.size(), f) != buf.size()) {
      std::cerr << "cannot write " << dump_path << "\n";
      if (f) fclose(f);
      return 1;
    }
    fclose(f);
  }
  m2_test_destroy(t);
  m2_close(ctx);
  return 0;
}
