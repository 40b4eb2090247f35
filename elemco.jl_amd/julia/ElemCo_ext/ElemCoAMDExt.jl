"""
    ElemCoAMDExt

Package extension of ElemCo.jl (v0.14.0) that runs the closed-shell coupled-cluster path
on an MI355X through `ElemCoAMD` (the `ccall` binding of `libelemco_hip.so`).  It goes in
ElemCo.jl's `ext/` directory next to `DmrgExt.jl`, with `ElemCoAMD` as a weak dependency:

    [weakdeps]
    ElemCoAMD = "1b757237-b3c0-5fc1-ab5e-6f510f10fec9"
    [extensions]
    ElemCoAMDExt = "ElemCoAMD"

and is selected like the DMRG extension (`Base.get_extension`, src/cc/ccdriver.jl:467-472)
from `eval_cc_groundstate` (src/cc/ccdriver.jl:348-393).  The functions below keep the
reference's signatures and side effects on `EC`:

- `calc_cc(EC, method)`                                  src/cc/cc.jl:2451-2516
- `calc_cc_resid(EC, T1, T2; dc, tworef, fixref, linearized)` src/cc/cc.jl:1343-1489
- `calc_lm_cc(EC, method)`                               src/cc/cc_lagrange.jl:1120-1223
- `calc_ccsd_vector_times_Jacobian(EC, U1, U2; dc)`      src/cc/cc_lagrange.jl:9-212
- `calc_1RDM(EC, U1, U2, T1, T2)`                        src/cc/cc_lagrange.jl:793-837
- `calc_pertT(EC, method; save_t3)`                      src/cc/cc_triples.jl:11-35

Amplitudes and multipliers go to the reference's scratch files (`T_vo`, `T_vvoo`,
`U_vo`, `U_vvoo`, through `save_current_singles` / `save_current_doubles`,
src/cc/cctools.jl:845-866), and `calc_pertT` / `calc_ccsd_vector_times_Jacobian` read them
back from there, as the reference does; on the device they stay resident between the
calls of one solve.  Not run in this repository (no Julia in its build image).
"""
module ElemCoAMDExt

using ElemCo
using ElemCo.Utils
using ElemCo.ECInfos
using ElemCo.ECMethods
using ElemCo.TensorTools
using ElemCo.FciDumps
using ElemCo.CCTools
using ElemCoAMD

# one device context per integral set: the in-memory dump of EC (EC.fd) is the key
const CONTEXTS = WeakKeyDict{Any,ElemCoAMD.ECGContext}()

"""
    gpu_context(EC::ECInfo)

The device context of `EC`'s integrals: the one-electron integrals and `(pq|rs)` of `EC.fd`
(restricted, closed-shell; `ST` of a similarity-transformed dump kept, so its exact DF
factors are two-sided), made once per dump.  The GPU path numbers orbitals as the
reference's default spaces do (occupied = the first nocc, setup_space_fd!,
src/infos/ecinfos.jl:171-192); other orbital strings are an error.
"""
function gpu_context(EC::ECInfo)
  get!(CONTEXTS, EC.fd) do
    EC.fd.uhf && error("the GPU path is closed-shell: unrestricted dumps are not supported")
    n = n_orbs(EC)
    nocc = n_occ_orbs(EC)
    (collect(EC.space['o']) == collect(1:nocc) && collect(EC.space['v']) == collect(nocc+1:n)) ||
      error("the GPU path needs occupied = 1:nocc and virtual = nocc+1:norb")
    h = ints1(EC, "::")
    eri = permutedims(ints2(EC, "::::"), (1, 3, 2, 4))     # <pr|qs> at [p,r,q,s] -> (pq|rs) at [p,q,r,s]
    st = is_similarity_transformed(EC.fd) ? 1 : 0
    ElemCoAMD.ECGContext(nocc, h, eri; st=st, e0=EC.fd.int0)
  end
end

_opts(EC::ECInfo) = ElemCoAMD.default_opts(thr=EC.options.cc.thr, conven=EC.options.cc.conven,
                                             maxit=EC.options.cc.maxit, shifts=EC.options.cc.shifts,
                                             shiftp=EC.options.cc.shiftp, maxdiis=EC.options.diis.maxdiis,
                                             resthr=EC.options.diis.resthr)

function _main_name(method::ECMethod)
  name = lowercase(method_name(method))
  (is_unrestricted(method) || has_prefix(method, "R")) && error("the GPU path is closed-shell only ($name)")
  return name
end

"`calc_cc(EC, method)` (src/cc/cc.jl:2451-2516) on the device; saves `T_vo`, `T_vvoo`."
function calc_cc(EC::ECInfo, method::ECMethod)
  ctx = gpu_context(EC)
  res = ElemCoAMD.calc_cc(ctx, _main_name(method); opts=_opts(EC))
  res["converged"] == 1.0 || println("WARNING: CC iterations did not converge!")
  save_current_singles(EC, ctx.T1)
  save_current_doubles(EC, ctx.T2)
  return OutDict("E" => res["E"], "ESS" => res["ESS"], "EOS" => res["EOS"])
end

"`calc_cc_resid(EC, T1, T2; dc)` (src/cc/cc.jl:1343-1489) on the device."
function calc_cc_resid(EC::ECInfo, T1, T2; dc=false, tworef=false, fixref=false, linearized=false)
  (tworef || fixref || linearized) && error("tworef / fixref / linearized residuals are not on the GPU path")
  return ElemCoAMD.calc_cc_resid(gpu_context(EC), Matrix{Float64}(T1), Array{Float64,4}(T2); dc)
end

"`calc_lm_cc(EC, method)` (src/cc/cc_lagrange.jl:1120-1223) on the device; saves `U_vo`, `U_vvoo`."
function calc_lm_cc(EC::ECInfo, method::ECMethod)
  ctx = gpu_context(EC)
  ctx.T1 = load2idx(EC, "T_vo")
  ctx.T2 = load4idx(EC, "T_vvoo")
  U1, U2 = ElemCoAMD.calc_lm_cc(ctx, _main_name(method); opts=_opts(EC))
  save_current_singles(EC, U1; prefix="U")
  save_current_doubles(EC, U2; prefix="U")
  return
end

"`calc_ccsd_vector_times_Jacobian(EC, U1, U2; dc)` (src/cc/cc_lagrange.jl:9-212) at `T_vo`, `T_vvoo`."
function calc_ccsd_vector_times_Jacobian(EC::ECInfo, U1, U2; dc=false)
  return ElemCoAMD.calc_ccsd_vector_times_Jacobian(gpu_context(EC), U1, U2; dc,
                                                   T1=load2idx(EC, "T_vo"), T2=load4idx(EC, "T_vvoo"))
end

"`calc_1RDM(EC, U1, U2, T1, T2)` (src/cc/cc_lagrange.jl:793-837) on the device: `(D1, dD1)`."
calc_1RDM(EC::ECInfo, U1, U2, T1, T2) = ElemCoAMD.calc_1RDM(gpu_context(EC), U1, U2, T1, T2)

"""
`calc_pertT(EC, method; save_t3)` (src/cc/cc_triples.jl:11-35), closed-shell branch: (T) or Λ-(T)
on `T_vo`, `T_vvoo` (and `U_vo`, `U_vvoo`).  Returns `OutDict("ET3" => E(T), "ET3b" => E[T])`.
"""
function calc_pertT(EC::ECInfo, method::ECMethod; save_t3=false)
  _main_name(method)
  save_t3 && error("saving perturbative triples is not supported on the GPU path")
  ctx = gpu_context(EC)
  T1, T2 = load2idx(EC, "T_vo"), load4idx(EC, "T_vvoo")
  if has_prefix(method, "Λ")
    et, etb = ElemCoAMD.pert_triples(ctx, T1, T2; U1=load2idx(EC, "U_vo"), U2=load4idx(EC, "U_vvoo"))
  else
    et, etb = ElemCoAMD.pert_triples(ctx, T1, T2)
  end
  return OutDict("ET3" => et, "ET3b" => etb)
end

end # module
