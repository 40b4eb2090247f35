"""
    ElemCoAMDChainRulesCoreExt

Reverse-mode rule of the CC correlation energy for ChainRulesCore-based automatic
differentiation (Zygote, Enzyme's rule import, ...).  The Λ equations are the adjoint of
the amplitude equations: at converged amplitudes T and multipliers Λ the CC Lagrangian
L(h, B; T, Λ) = E(T) + Λ·R(T) is stationary in T and Λ, so dE/dh = ∂L/∂h and
dE/dB = ∂L/∂B, which the library evaluates by walking the residual's contractions in
reverse on the device (`ecg_energy_grad`).  No v⁴-sized intermediate is stored: the
particle ladder's adjoint is rebuilt from B tiles block by block.

Not run in this repository (no Julia in its build image); the gradient itself is tested
through the library's C ABI against central finite differences of re-solved energies.
"""
module ElemCoAMDChainRulesCoreExt

using ElemCoAMD
using ChainRulesCore

function ChainRulesCore.rrule(::typeof(ElemCoAMD.cc_energy), h::Matrix{Float64}, BL::Array{Float64,3},
                              nocc::Integer; BR=nothing, e0=0.0, method="ccsd", thr=1e-14, device=0)
  BR === nothing || error("the rrule of cc_energy takes Hermitian integrals (BR = nothing)")
  ctx = ElemCoAMD.ECGContext(nocc, h, BL; e0, device)
  opts = ElemCoAMD.default_opts(thr=thr, maxit=200)
  res = ElemCoAMD.calc_cc(ctx, method; opts)
  res["converged"] == 1.0 || error("cc_energy: CC iterations did not converge")
  ElemCoAMD.calc_lm_cc(ctx, method; opts)
  gh, gb = ElemCoAMD.energy_grad(ctx, method)
  function cc_energy_pullback(Ebar)
    Ē = unthunk(Ebar)
    return (NoTangent(), Ē .* gh, Ē .* gb, NoTangent())
  end
  return res["E"], cc_energy_pullback
end

end # module
