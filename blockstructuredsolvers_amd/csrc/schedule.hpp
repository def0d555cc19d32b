// schedule.hpp -- recursion-level schedules (host only).
//
// A level of N blocks is split into chains of consecutive blocks separated by single
// separator blocks; the separators form the next level's (reduced) system.  The
// last level is one chain with no separators, solved by the same chain kernels.
#pragma once

#include <algorithm>
#include <array>
#include <cstdint>
#include <vector>

namespace bss {

struct Chain {
  int start, len, sepL, sepR;  // separators are indices into this level's separator list
};

struct LevelShape {
  int N = 0;                     // blocks at this level
  std::vector<Chain> chains;     // non-empty chains only
  std::vector<int> seppos;       // block index of each separator (P = size)
};

// The reference's _build_sequence (src/gpu_batched.jl:112-143): from the top level
// walk down, choosing the smallest m >= 3 with (N + m) % (m + 1) == 0 and P >= 3,
// preferring m < P; returns (N_i, m_i, P_i) top level first (the reference builds the
// lists bottom first with pushfirst!).  Empty when no level exists.
inline std::vector<std::array<int, 3>> reference_levels(int start) {
  std::vector<std::array<int, 3>> out;
  long long ncur = start;
  while (true) {
    int best_m = -1, best_P = -1, backup_m = -1, backup_P = -1;
    const long long max_m = (ncur - 3) / 2;
    for (long long m = 3; m <= max_m; ++m) {
      const long long num = ncur + m, den = m + 1;
      if (num % den == 0) {
        const long long pprev = num / den;
        if (pprev >= 3) {
          if (m < pprev) {
            best_m = (int)m;
            best_P = (int)pprev;
            break;
          }
          if (backup_m < 0) {
            backup_m = (int)m;
            backup_P = (int)pprev;
          }
        }
      }
    }
    if (best_m < 0 && backup_m >= 0) {
      best_m = backup_m;
      best_P = backup_P;
    }
    if (best_m < 0) break;
    out.push_back({(int)ncur, best_m, best_P});
    ncur = best_P;
  }
  return out;
}

// find_sequence_upper (src/gpu_batched.jl:155-162): smallest N_use >= N with a
// schedule.  The reference returns only the lists; N_use is their top level.
inline int reference_schedule(int N, std::vector<std::array<int, 3>>* levels) {
  int n_use = std::max(N, 1);
  while (true) {
    auto lv = reference_levels(n_use);
    if (!lv.empty()) {
      if (levels) *levels = lv;
      return n_use;
    }
    ++n_use;
  }
}

// Level with separators at 0-based blocks k(m+1), k = 0..P-1 (the reference's layout,
// src/gpu_batched.jl:211-215): chains of m blocks between consecutive separators.
inline LevelShape reference_level(int N, int m, int P) {
  LevelShape L;
  L.N = N;
  for (int k = 0; k < P; ++k) L.seppos.push_back(k * (m + 1));
  for (int k = 0; k + 1 < P; ++k) L.chains.push_back({k * (m + 1) + 1, m, k, k + 1});
  return L;
}

// Single chain over the whole level: the recursion base.
inline LevelShape base_level(int N) {
  LevelShape L;
  L.N = N;
  L.chains.push_back({0, N, -1, -1});
  return L;
}

// This library's schedule: Q chains whose lengths differ by at most one, the first and
// last chains touching the ends of the level (so neither end is a separator and the
// first chain carries no panel).  Q = round((N+1)/(m+1)); every chain has >= 1 block.
inline LevelShape split_level(int N, int m) {
  LevelShape L;
  L.N = N;
  int Q = (int)((N + 1 + (m + 1) / 2) / (m + 1));
  Q = std::max(2, std::min(Q, (N + 1) / 2));
  const int P = Q - 1;
  const int interior = N - P;
  const int base = interior / Q, extra = interior % Q;
  int pos = 0;
  for (int c = 0; c < Q; ++c) {
    const int len = base + (c < extra ? 1 : 0);
    L.chains.push_back({pos, len, c > 0 ? c - 1 : -1, c < P ? c : -1});
    pos += len;
    if (c < P) {
      L.seppos.push_back(pos);
      ++pos;
    }
  }
  return L;
}

// Full schedule: levels top first, the last one a single chain.
// chain_len[k] > 0 fixes level k's chain length; other levels use `m_default`.
inline std::vector<LevelShape> auto_schedule(int N, const std::vector<int>& chain_len,
                                             int m_default, int base_max) {
  std::vector<LevelShape> out;
  int cur = N;
  size_t li = 0;
  while (cur > base_max && cur >= 3) {
    const int m = (li < chain_len.size() && chain_len[li] > 0) ? chain_len[li] : m_default;
    LevelShape L = split_level(cur, std::max(1, m));
    const int P = (int)L.seppos.size();
    out.push_back(std::move(L));
    cur = P;
    ++li;
  }
  out.push_back(base_level(cur));
  return out;
}

inline std::vector<LevelShape> reference_schedule_levels(int N, int* n_use) {
  std::vector<std::array<int, 3>> lv;
  *n_use = reference_schedule(N, &lv);
  std::vector<LevelShape> out;
  for (const auto& t : lv) out.push_back(reference_level(t[0], t[1], t[2]));
  out.push_back(base_level(lv.back()[2]));
  return out;
}

}  // namespace bss
