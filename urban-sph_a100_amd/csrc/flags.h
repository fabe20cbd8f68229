// Neighbour-list entry layout shared by the search (nnps.hip), the physics
// passes and the PPE.  Each entry is a 32-bit word: the neighbour index in
// the low 26 bits and six selection flags, each the exact FP64 evaluation of
// one kernel's window and radius rule at search time:
//   F_DP  dcell window, r + 1e-20 < kappa*max(hi,hj)  (advection, PPE)
//   F_DM  dcell window, r - 1e-20 < kappa*max(hi,hj)  (pressure force)
//   F_HP  dcell window, r + 1e-20 < kappa*hi          (prep / LES)
//   F_HM  dcell window, r - 1e-20 < kappa*hi          (wall velocity, free-slip)
//   F_B   binning window, 0 < r < kappa*hi           (gradient correction)
//   F_BH  binning window, r + 1e-20 < kappa*(0.5*hi) (IBM interpolation)
#pragma once

namespace sph {
constexpr unsigned NB_IDX_BITS = 26;
constexpr unsigned NB_IDX = (1u << NB_IDX_BITS) - 1;
constexpr unsigned F_DP = 1u << 26;
constexpr unsigned F_DM = 1u << 27;
constexpr unsigned F_HP = 1u << 28;
constexpr unsigned F_HM = 1u << 29;
constexpr unsigned F_B = 1u << 30;
constexpr unsigned F_BH = 1u << 31;
constexpr int MAX_PARTICLES = (1 << NB_IDX_BITS) - 1;
}  // namespace sph
