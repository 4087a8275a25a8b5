// abi_layout.cpp -- compiled against include/tm/tm.h with the host C++ compiler
// (no HIP headers): fixes the struct layouts the ctypes binding of
// trophic_maelstorm_amd mirrors, and prints them (struct, field, offset, size)
// for the test that compares the two.
#include <cstddef>
#include <cstdio>
#include "tm/tm.h"

// The layouts the binding relies on: 8-byte members first, the two int32 tails
// packed into one 8-byte slot, no padding anywhere.
static_assert(sizeof(tm_dp_params) == 88 && offsetof(tm_dp_params, scheme) == 80, "tm_dp_params");
static_assert(sizeof(tm_rietkerk_params) == 144 && offsetof(tm_rietkerk_params, sampler) == 136,
              "tm_rietkerk_params");
static_assert(sizeof(tm_instance_desc) == 24 && offsetof(tm_instance_desc, step0) == 16, "tm_instance_desc");
static_assert(offsetof(tm_dpc_params, A) == 136 && offsetof(tm_dpc_params, sampler) == 352 &&
                  sizeof(tm_dpc_params) == 360,
              "tm_dpc_params");
static_assert(sizeof(tm_fke_move_species) == 16, "tm_fke_move_species");
static_assert(offsetof(tm_fke_params, D_P) == 264 && offsetof(tm_fke_params, mv_invariance) == 440 &&
                  sizeof(tm_fke_params) == 456,
              "tm_fke_params");

#define F(S, m) std::printf("%s %s %zu %zu\n", #S, #m, offsetof(S, m), sizeof(((S*)0)->m))
#define SZ(S) std::printf("%s sizeof %zu %zu\n", #S, sizeof(S), sizeof(S))

int main()
{
    SZ(tm_dp_params);
    F(tm_dp_params, a); F(tm_dp_params, b); F(tm_dp_params, D); F(tm_dp_params, sigma); F(tm_dp_params, dt);
    F(tm_dp_params, dx); F(tm_dp_params, a_end); F(tm_dp_params, ramp_steps); F(tm_dp_params, flush_eps);
    F(tm_dp_params, flush_neg); F(tm_dp_params, scheme); F(tm_dp_params, sampler);
    SZ(tm_rietkerk_params);
    F(tm_rietkerk_params, R); F(tm_rietkerk_params, c); F(tm_rietkerk_params, gmax); F(tm_rietkerk_params, alpha);
    F(tm_rietkerk_params, rW); F(tm_rietkerk_params, W0); F(tm_rietkerk_params, d); F(tm_rietkerk_params, D);
    F(tm_rietkerk_params, K); F(tm_rietkerk_params, sigma); F(tm_rietkerk_params, dt); F(tm_rietkerk_params, dx);
    F(tm_rietkerk_params, flush_eps); F(tm_rietkerk_params, sampler); F(tm_rietkerk_params, reserved);
    SZ(tm_dpc_params);
    F(tm_dpc_params, a); F(tm_dpc_params, b); F(tm_dpc_params, D); F(tm_dpc_params, sigma); F(tm_dpc_params, v);
    F(tm_dpc_params, M); F(tm_dpc_params, E); F(tm_dpc_params, A); F(tm_dpc_params, H); F(tm_dpc_params, pR);
    F(tm_dpc_params, dt); F(tm_dpc_params, dx); F(tm_dpc_params, a_end); F(tm_dpc_params, ramp_steps);
    F(tm_dpc_params, flush_eps); F(tm_dpc_params, flush_neg); F(tm_dpc_params, sampler);
    F(tm_dpc_params, reserved);
    SZ(tm_instance_desc);
    F(tm_instance_desc, L); F(tm_instance_desc, params_index); F(tm_instance_desc, seed);
    F(tm_instance_desc, instance_id); F(tm_instance_desc, step0);
    SZ(tm_fke_move_species);
    F(tm_fke_move_species, rD); F(tm_fke_move_species, rV); F(tm_fke_move_species, mu_prefactor);
    SZ(tm_fke_rk_params);
    F(tm_fke_rk_params, a); F(tm_fke_rk_params, c); F(tm_fke_rk_params, gmax); F(tm_fke_rk_params, alpha);
    F(tm_fke_rk_params, rW); F(tm_fke_rk_params, W0); F(tm_fke_rk_params, K); F(tm_fke_rk_params, A);
    F(tm_fke_rk_params, H); F(tm_fke_rk_params, E); F(tm_fke_rk_params, M); F(tm_fke_rk_params, D_W);
    F(tm_fke_rk_params, D_O); F(tm_fke_rk_params, dt); F(tm_fke_rk_params, dt_rk); F(tm_fke_rk_params, dx);
    SZ(tm_fke_params);
    F(tm_fke_params, a); F(tm_fke_params, c); F(tm_fke_params, gmax); F(tm_fke_params, alpha); F(tm_fke_params, rW);
    F(tm_fke_params, W0); F(tm_fke_params, K); F(tm_fke_params, A); F(tm_fke_params, H); F(tm_fke_params, E);
    F(tm_fke_params, M); F(tm_fke_params, D_P); F(tm_fke_params, D_W); F(tm_fke_params, D_O);
    F(tm_fke_params, D_mov); F(tm_fke_params, v); F(tm_fke_params, sigma); F(tm_fke_params, pR);
    F(tm_fke_params, dt); F(tm_fke_params, dt_analytic); F(tm_fke_params, dt_adv); F(tm_fke_params, dx);
    F(tm_fke_params, flush_eps); F(tm_fke_params, mv_invariance); F(tm_fke_params, CI); F(tm_fke_params, sampler);
    F(tm_fke_params, reserved);
    return 0;
}
