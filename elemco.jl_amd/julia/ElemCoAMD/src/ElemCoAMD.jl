"""
    ElemCoAMD

Julia binding of `libelemco_hip.so`, the MI355X (gfx950) library of the
closed-shell coupled-cluster path of ElemCo.jl.  Every wrapper is a `ccall` of
one entry point of `include/elemco_hip.h`; arrays are passed as Julia
column-major `Float64` arrays (the layout the library expects), and a negative
status is raised with `error(...)`, the reference's convention (e.g.
`src/cc/cc.jl:1364`).

Context-level methods (an `ECGContext` stands for the reference's `EC::ECInfo`
on this path: it owns the integrals resident on one GPU and, like the
reference's scratch files `T_vo`, `T_vvoo`, `U_vo`, `U_vvoo`, the amplitudes and
multipliers of the last solves):

- `calc_cc(ctx, method)`                               ↔ `src/cc/cc.jl:2451-2516`
- `calc_cc_resid(ctx, T1, T2; dc)`                     ↔ `src/cc/cc.jl:1343-1489`
- `calc_lm_cc(ctx, method)`                            ↔ `src/cc/cc_lagrange.jl:1120-1223`
- `calc_ccsd_vector_times_Jacobian(ctx, U1, U2; dc)`   ↔ `src/cc/cc_lagrange.jl:9-212` (T from the context)
- `calc_1RDM(ctx, U1, U2, T1, T2)`                     ↔ `src/cc/cc_lagrange.jl:793-837`
- `calc_pertT(ctx, method; save_t3)`                   ↔ `src/cc/cc_triples.jl:11-35` (T, U from the context)
- `cc_energy(h, BL, nocc; ...)`, differentiable through the `ChainRulesCore.rrule`
  of the package extension `ext/ElemCoAMDChainRulesCoreExt.jl` (`energy_grad`).

The methods with the reference's own `EC::ECInfo` signatures live in the
ElemCo.jl package extension `julia/ElemCo_ext/ElemCoAMDExt.jl`, which builds the
context from `EC.fd` and reads and writes the reference's scratch files.  Build a
context from MO density-fitting factors `B[p,q,Q]` (the layout of `mmL`,
`src/integrals/dftools.jl:102`) and the one-electron integrals, from an FCIDUMP
file (`ECGContext(path)`), or from four-index integrals in memory
(`ECGContext(nocc, h, eri; ...)`).
"""
module ElemCoAMD

export ECGContext, write_fcidump, df_freeze, calc_cc_resid, calc_ccsd_vector_times_Jacobian, calc_1RDM, calc_pertT, calc_cc, calc_lm_cc,
       energy_grad, cc_energy, pert_triples, SolverOpts, default_opts, ECG_FP64, ECG_FP32, ECG_BF16X3

# the library sits in elemco.jl_amd/, three levels above this file (julia/ElemCoAMD/src)
const LIB = get(ENV, "ELEMCO_HIP_LIB", joinpath(@__DIR__, "..", "..", "..", "libelemco_hip.so"))

const ECG_CCSD = Cint(0)
const ECG_DCSD = Cint(1)
const ECG_FP64 = Cint(0)
const ECG_FP32 = Cint(1)
const ECG_BF16X3 = Cint(2)

"Mirror of `ecg_solver_opts` (defaults of `src/infos/options.jl:128-221,272-282`)."
mutable struct SolverOpts
  thr::Cdouble
  conven::Cdouble
  shifts::Cdouble
  shiftp::Cdouble
  resthr::Cdouble
  maxit::Cint
  maxdiis::Cint
  refine_every::Cint
  graphs::Cint
end

function default_opts(; kw...)
  o = SolverOpts(0, 0, 0, 0, 0, 0, 0, 0, 0)
  ccall((:ecg_default_opts, LIB), Cvoid, (Ref{SolverOpts},), o)
  for (k, v) in kw
    setfield!(o, k, convert(fieldtype(SolverOpts, k), v))
  end
  return o
end

mutable struct ECGContext
  handle::Ptr{Cvoid}
  nocc::Int
  nvirt::Int
  naux::Int
  # amplitudes (T) and contravariant multipliers (U) of the last calc_cc / calc_lm_cc:
  # the reference keeps them in the scratch files T_vo, T_vvoo, U_vo, U_vvoo
  T1::Matrix{Float64}
  T2::Array{Float64,4}
  U1::Matrix{Float64}
  U2::Array{Float64,4}
  method::String
end

ECGContext(handle::Ptr{Cvoid}, nocc::Integer, nvirt::Integer, naux::Integer) =
  ECGContext(handle, nocc, nvirt, naux, zeros(0, 0), zeros(0, 0, 0, 0), zeros(0, 0), zeros(0, 0, 0, 0), "")

function _check(ctx, st)
  if st != 0
    msg = unsafe_string(ccall((:ecg_last_error, LIB), Cstring, (Ptr{Cvoid},), ctx === nothing ? C_NULL : ctx.handle))
    error(msg)
  end
  return nothing
end

"""
    ECGContext(nocc, h, BL; BR=nothing, e0=0.0, device=0)

Upload the one-electron integrals `h` (n×n) and DF factors `BL` (n×n×naux;
`BR` for non-Hermitian integrals) to HIP device `device` (`ecg_create`).
"""
function ECGContext(nocc::Integer, h::Matrix{Float64}, BL::Array{Float64,3}; BR=nothing, e0=0.0, device=0)
  n = size(h, 1)
  naux = size(BL, 3)
  size(BL, 1) == n && size(BL, 2) == n || error("BL must be n×n×naux")
  ref = Ref{Ptr{Cvoid}}(C_NULL)
  brp = BR === nothing ? Ptr{Cdouble}(C_NULL) : pointer(BR)
  st = GC.@preserve BR ccall((:ecg_create, LIB), Cint,
                             (Ref{Ptr{Cvoid}}, Cint, Cint, Cint, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Cdouble, Cint,
                              Csize_t),
                             ref, nocc, n - nocc, naux, h, BL, brp, e0, device, 0)
  _check(nothing, st)
  ctx = ECGContext(ref[], nocc, n - nocc, naux)
  finalizer(c -> (c.handle != C_NULL && ccall((:ecg_destroy, LIB), Cint, (Ptr{Cvoid},), c.handle); c.handle = C_NULL), ctx)
  return ctx
end

"""
    ECGContext(fcidump::AbstractString; rtol=0.0, device=0, freeze_orbs=Int[])

`read_fcidump` (src/integrals/dump.jl:326-346; text or NPY2/NPY1 integrals, plain or
gzip-compressed), exact DF factors and `ecg_create` with nocc = NELEC/2 in one native
call (`ecg_create_from_fcidump`): the set-up `ccdriver(EC, method; fcidump)` does
(src/cc/ccdriver.jl:37-76) for the closed-shell path.  A non-empty `freeze_orbs`
(1-based) applies `@freeze_orbs` first (`freeze_orbs_in_dump`,
src/integrals/dumptools.jl:20-62; `ecg_fcidump_freeze`).
"""
function ECGContext(fcidump::AbstractString; rtol=0.0, device=0, freeze_orbs=Int[])
  ref = Ref{Ptr{Cvoid}}(C_NULL)
  if isempty(freeze_orbs)
    st = ccall((:ecg_create_from_fcidump, LIB), Cint, (Ref{Ptr{Cvoid}}, Cstring, Cdouble, Cint, Csize_t),
               ref, fcidump, rtol, device, 0)
  else
    fd = Ref{Ptr{Cvoid}}(C_NULL)
    _check(nothing, ccall((:ecg_fcidump_read, LIB), Cint, (Cstring, Ref{Ptr{Cvoid}}), fcidump, fd))
    orbs = Cint.(collect(freeze_orbs))
    st = ccall((:ecg_fcidump_freeze, LIB), Cint, (Ptr{Cvoid}, Ptr{Cint}, Cint, Ptr{Cint}, Ptr{Cint}),
               fd[], orbs, length(orbs), C_NULL, C_NULL)
    if st == 0
      st = ccall((:ecg_create_from_fcidump_handle, LIB), Cint,
                 (Ref{Ptr{Cvoid}}, Ptr{Cvoid}, Cdouble, Cint, Csize_t), ref, fd[], rtol, device, 0)
    end
    ccall((:ecg_fcidump_free, LIB), Cint, (Ptr{Cvoid},), fd[])
  end
  _check(nothing, st)
  o, v, q = Ref{Cint}(0), Ref{Cint}(0), Ref{Cint}(0)
  ccall((:ecg_dims, LIB), Cint, (Ptr{Cvoid}, Ref{Cint}, Ref{Cint}, Ref{Cint}), ref[], o, v, q)
  ctx = ECGContext(ref[], o[], v[], q[])
  finalizer(c -> (c.handle != C_NULL && ccall((:ecg_destroy, LIB), Cint, (Ptr{Cvoid},), c.handle); c.handle = C_NULL), ctx)
  return ctx
end

"""
    ECGContext(nocc, C, h_ao, AAP, PQ; tol=1e-8, e0=0.0, device=0)

DF set-up from AO integrals (`ecg_create_from_ao`): `B[p,q,L]` of
`generate_3idx_integrals` (src/integrals/dftools.jl:92-125) with `sqrtinvchol` of the
metric `PQ` (src/tools/tensortools.jl:377-391), `h = C' h_ao C`; `C` is nao×nmo, `AAP`
nao×nao×nfit (the reference's `AAP` before the metric), `PQ` nfit×nfit.
"""
function ECGContext(nocc::Integer, C::Matrix{Float64}, h_ao::Matrix{Float64}, AAP::Array{Float64,3},
                    PQ::Matrix{Float64}; tol=1e-8, e0=0.0, device=0)
  nao, nmo = size(C)
  nfit = size(PQ, 1)
  ref = Ref{Ptr{Cvoid}}(C_NULL)
  st = ccall((:ecg_create_from_ao, LIB), Cint,
             (Ref{Ptr{Cvoid}}, Cint, Ptr{Cdouble}, Cint, Cint, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Cint,
              Cdouble, Cdouble, Cint, Csize_t),
             ref, nocc, C, nao, nmo, h_ao, AAP, PQ, nfit, tol, e0, device, 0)
  _check(nothing, st)
  o, v, q = Ref{Cint}(0), Ref{Cint}(0), Ref{Cint}(0)
  ccall((:ecg_dims, LIB), Cint, (Ptr{Cvoid}, Ref{Cint}, Ref{Cint}, Ref{Cint}), ref[], o, v, q)
  ctx = ECGContext(ref[], o[], v[], q[])
  finalizer(c -> (c.handle != C_NULL && ccall((:ecg_destroy, LIB), Cint, (Ptr{Cvoid},), c.handle); c.handle = C_NULL), ctx)
  return ctx
end

"""
    ECGContext(nocc, h, eri::Array{Float64,4}; st=0, e0=0.0, rtol=0.0, device=0)

Context from four-index integrals in memory, `eri[p,q,r,s] = (pq|rs)` (chemists' notation,
n^4) and `h` (n×n), e.g. the arrays of the reference's in-memory `FDump`
(src/integrals/dump.jl:107-128): `ecg_fcidump_from_integrals` makes a dump
(NORB = n, NELEC = 2 nocc, MS2 = 0, ST = `st`), then its exact DF factors and the
context (`ecg_create_from_fcidump_handle`).
"""
function ECGContext(nocc::Integer, h::Matrix{Float64}, eri::Array{Float64,4}; st=0, e0=0.0, rtol=0.0, device=0)
  n = size(h, 1)
  size(eri) == (n, n, n, n) || error("eri must be n×n×n×n")
  fd = Ref{Ptr{Cvoid}}(C_NULL)
  _check(nothing, ccall((:ecg_fcidump_from_integrals, LIB), Cint,
                        (Cint, Cint, Cint, Cint, Ptr{Cdouble}, Ptr{Cdouble}, Cdouble, Ref{Ptr{Cvoid}}),
                        n, 2 * nocc, 0, st, h, eri, e0, fd))
  ref = Ref{Ptr{Cvoid}}(C_NULL)
  st2 = ccall((:ecg_create_from_fcidump_handle, LIB), Cint, (Ref{Ptr{Cvoid}}, Ptr{Cvoid}, Cdouble, Cint, Csize_t),
              ref, fd[], rtol, device, 0)
  ccall((:ecg_fcidump_free, LIB), Cint, (Ptr{Cvoid},), fd[])
  _check(nothing, st2)
  o, v, q = Ref{Cint}(0), Ref{Cint}(0), Ref{Cint}(0)
  ccall((:ecg_dims, LIB), Cint, (Ptr{Cvoid}, Ref{Cint}, Ref{Cint}, Ref{Cint}), ref[], o, v, q)
  ctx = ECGContext(ref[], o[], v[], q[])
  finalizer(c -> (c.handle != C_NULL && ccall((:ecg_destroy, LIB), Cint, (Ptr{Cvoid},), c.handle); c.handle = C_NULL), ctx)
  return ctx
end

_z1(ctx) = zeros(ctx.nvirt, ctx.nocc)
_z2(ctx) = zeros(ctx.nvirt, ctx.nvirt, ctx.nocc, ctx.nocc)

# method strings of the closed-shell path ("ccsd", "dcsd", "ccsd(t)", "λccsd(t)", "Λccsd", ...):
# (main method code, Λ prefix, perturbative triples)
function _parse_method(method::AbstractString)
  m = lowercase(strip(method))
  lam = startswith(m, "λ")
  lam && (m = m[nextind(m, 1):end])
  pert = endswith(m, "(t)")
  pert && (m = m[1:end-3])
  m in ("ccsd", "dcsd") || error("method $(method) is not on the closed-shell GPU path (ccsd, dcsd, [λ]ccsd(t))")
  return (m == "dcsd" ? ECG_DCSD : ECG_CCSD), lam, pert
end

"`calc_cc_resid(EC, T1, T2; dc)` (src/cc/cc.jl:1343-1489): returns freshly allocated `(R1, R2)`."
function calc_cc_resid(ctx::ECGContext, T1::Matrix{Float64}, T2::Array{Float64,4}; dc=false, precision=ECG_FP64)
  R1, R2 = _z1(ctx), _z2(ctx)
  st = ccall((:ecg_cc_residual, LIB), Cint,
             (Ptr{Cvoid}, Cint, Cint, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}),
             ctx.handle, dc ? ECG_DCSD : ECG_CCSD, precision, T1, T2, R1, R2)
  _check(ctx, st)
  return R1, R2
end

"""
`calc_ccsd_vector_times_Jacobian(EC, U1, U2; dc)` (src/cc/cc_lagrange.jl:9-212): the Λ residual
at the amplitudes of the context's last `calc_cc` (the reference reads them from `T_vo` /
`T_vvoo`); `T1`, `T2` override them.
"""
function calc_ccsd_vector_times_Jacobian(ctx::ECGContext, U1, U2; dc=false, precision=ECG_FP64, T1=ctx.T1, T2=ctx.T2)
  isempty(T1) && error("no amplitudes in the context: run calc_cc first")
  R1, R2 = _z1(ctx), _z2(ctx)
  st = ccall((:ecg_lambda_residual, LIB), Cint,
             (Ptr{Cvoid}, Cint, Cint, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble},
              Ptr{Cdouble}),
             ctx.handle, dc ? ECG_DCSD : ECG_CCSD, precision, T1, T2, U1, U2, R1, R2)
  _check(ctx, st)
  return R1, R2
end

"`calc_1RDM(EC, U1, U2, T1, T2)` (src/cc/cc_lagrange.jl:793-837): returns `(D1, dD1)`."
function calc_1RDM(ctx::ECGContext, U1, U2, T1, T2)
  n = ctx.nocc + ctx.nvirt
  D1, dD1 = zeros(n, n), zeros(n, n)
  st = ccall((:ecg_one_rdm, LIB), Cint,
             (Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}),
             ctx.handle, T1, T2, U1, U2, D1, dD1)
  _check(ctx, st)
  return D1, dD1
end

"""
    pert_triples(ctx, T1, T2; U1=nothing, U2=nothing, precision=ECG_FP64, first=0, count=-1)

(T) of `calc_pertT_closed_shell` (src/cc/cc_triples.jl:39-148), Λ-(T) of
`calc_ΛpertT_closed_shell` (:159-288) when `U1`, `U2` are given, on explicit amplitudes:
`(E(T), E[T])`.  `first`/`count` restrict it to a range of triplets of the reference's
loop order (partial sums add up).
"""
function pert_triples(ctx::ECGContext, T1, T2; U1=nothing, U2=nothing, precision=ECG_FP64, first=0, count=-1)
  lam = U1 !== nothing
  et, etb = Ref{Cdouble}(0), Ref{Cdouble}(0)
  u1 = lam ? U1 : Float64[]
  u2 = lam ? U2 : Float64[]
  st = if count < 0
    ccall((:ecg_pert_triples, LIB), Cint,
          (Ptr{Cvoid}, Cint, Cint, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ref{Cdouble}, Ref{Cdouble}),
          ctx.handle, lam, precision, T1, T2, u1, u2, et, etb)
  else
    ccall((:ecg_pert_triples_range, LIB), Cint,
          (Ptr{Cvoid}, Cint, Cint, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Clong, Clong, Ref{Cdouble},
           Ref{Cdouble}),
          ctx.handle, lam, precision, T1, T2, u1, u2, first, count, et, etb)
  end
  _check(ctx, st)
  return et[], etb[]
end

"""
    calc_pertT(ctx, method; save_t3=false, precision=ECG_FP64)

`calc_pertT(EC, method; save_t3)` (src/cc/cc_triples.jl:11-35), closed-shell branch: (T) for
"ccsd(t)", Λ-(T) for "λccsd(t)", on the amplitudes (and multipliers) of the context's last
`calc_cc` (and `calc_lm_cc`), as the reference reads `T_vo`, `T_vvoo` (`U_vo`, `U_vvoo`).
Returns `Dict("ET3" => E(T), "ET3b" => E[T])`.  `save_t3` is not supported (an error, as the
reference's assertion for the Λ branch).
"""
function calc_pertT(ctx::ECGContext, method::AbstractString; save_t3=false, precision=ECG_FP64)
  save_t3 && error("saving perturbative triples is not supported on the GPU path")
  _, lam, _ = _parse_method(method)
  isempty(ctx.T1) && error("no amplitudes in the context: run calc_cc first")
  lam && isempty(ctx.U1) && error("no Λ multipliers in the context: run calc_lm_cc first")
  et, etb = lam ? pert_triples(ctx, ctx.T1, ctx.T2; U1=ctx.U1, U2=ctx.U2, precision) :
                  pert_triples(ctx, ctx.T1, ctx.T2; precision)
  return Dict("ET3" => et, "ET3b" => etb)
end

"""
`calc_cc(EC, method)` (src/cc/cc.jl:2451-2516) for "ccsd" / "dcsd" (a Λ prefix or a "(t)" suffix
is ignored here, as `eval_cc_groundstate` passes the main method): iterate on the device from
the MP2 guess; the converged amplitudes are kept in the context (the reference's `T_vo`,
`T_vvoo`).  Returns `Dict("E" => E_H, "ESS", "EOS", "Eproj", "dE", "NormR", "converged")`.
"""
function calc_cc(ctx::ECGContext, method::AbstractString; opts=default_opts(), precision=ECG_FP64)
  m, _, _ = _parse_method(method)
  T1, T2 = _z1(ctx), _z2(ctx)
  e = zeros(6)
  it, conv = Ref{Cint}(0), Ref{Cint}(0)
  st = ccall((:ecg_cc_solve, LIB), Cint,
             (Ptr{Cvoid}, Cint, Cint, Ref{SolverOpts}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ref{Cint}, Ref{Cint}),
             ctx.handle, m, precision, opts, T1, T2, e, it, conv)
  _check(ctx, st)
  conv[] == 0 && @warn "CC iterations did not converge in $(it[]) iterations"
  ctx.T1, ctx.T2, ctx.method = T1, T2, lowercase(method)
  ctx.U1, ctx.U2 = zeros(0, 0), zeros(0, 0, 0, 0)
  return Dict("E" => e[1], "ESS" => e[2], "EOS" => e[3], "Eproj" => e[4], "dE" => e[5], "NormR" => e[6],
              "converged" => Float64(conv[]))
end

"""
`calc_lm_cc(EC, method)` (src/cc/cc_lagrange.jl:1120-1223): Λ multipliers at the amplitudes of the
context's last `calc_cc`, kept in the context (the reference's `U_vo`, `U_vvoo`) and returned
as `(U1, U2)`.
"""
function calc_lm_cc(ctx::ECGContext, method::AbstractString; opts=default_opts(), precision=ECG_FP64)
  m, _, _ = _parse_method(method)
  isempty(ctx.T1) && error("no amplitudes in the context: run calc_cc first")
  U1, U2 = _z1(ctx), _z2(ctx)
  it, conv = Ref{Cint}(0), Ref{Cint}(0)
  st = ccall((:ecg_lambda_solve, LIB), Cint,
             (Ptr{Cvoid}, Cint, Cint, Ref{SolverOpts}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble},
              Ref{Cint}, Ref{Cint}),
             ctx.handle, m, precision, opts, ctx.T1, ctx.T2, U1, U2, it, conv)
  _check(ctx, st)
  conv[] == 0 && @warn "Lambda iterations did not converge in $(it[]) iterations"
  ctx.U1, ctx.U2 = U1, U2
  return U1, U2
end

"""
`energy_grad(ctx, method)`: `(dE/dh [n,n], dE/dB [n,n,naux])` of the correlation energy at the
context's converged amplitudes and Λ multipliers (`ecg_energy_grad`; no reference
counterpart: the backward the ChainRulesCore `rrule` of `cc_energy` calls).
"""
function energy_grad(ctx::ECGContext, method::AbstractString=ctx.method)
  m, _, _ = _parse_method(method)
  (isempty(ctx.T1) || isempty(ctx.U1)) && error("run calc_cc and calc_lm_cc first")
  n = ctx.nocc + ctx.nvirt
  gh, gb = zeros(n, n), zeros(n, n, ctx.naux)
  st = ccall((:ecg_energy_grad, LIB), Cint,
             (Ptr{Cvoid}, Cint, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}),
             ctx.handle, m, ctx.T1, ctx.T2, ctx.U1, ctx.U2, gh, gb)
  _check(ctx, st)
  return gh, gb
end

"""
    cc_energy(h, BL, nocc; BR=nothing, e0=0.0, method="ccsd", thr=1e-14, device=0)

CCSD / DCSD correlation energy of the integrals (h, BL[, BR]) with `nocc` doubly occupied
orbitals: the primal of the ChainRulesCore `rrule` of the package extension
`ElemCoAMDChainRulesCoreExt`, whose pullback returns (dE/dh, dE/dBL[, dE/dBR]) from the Λ
equations (`energy_grad`).  The solve is tight (`thr`) because the gradient assumes
stationarity in T and Λ.
"""
function cc_energy(h::Matrix{Float64}, BL::Array{Float64,3}, nocc::Integer; BR=nothing, e0=0.0, method="ccsd",
                   thr=1e-14, device=0)
  ctx = ECGContext(nocc, h, BL; BR, e0, device)
  res = calc_cc(ctx, method; opts=default_opts(thr=thr, maxit=200))
  return res["E"]
end

"""
    write_fcidump(fcidump_in, fcidump_out; freeze_orbs=Int[], tol=1e-12)

`@freeze_orbs` (freeze_orbs_in_dump, src/integrals/dumptools.jl:20-62) then `@write_ints`
(write_fcidump, src/integrals/dump.jl:752-945) on a dump read by the library's reader
(`ecg_fcidump_read`, `ecg_fcidump_freeze`, `ecg_fcidump_write`).
"""
function write_fcidump(fcidump_in::AbstractString, fcidump_out::AbstractString; freeze_orbs=Int[], tol=1e-12)
  fd = Ref{Ptr{Cvoid}}(C_NULL)
  _check(nothing, ccall((:ecg_fcidump_read, LIB), Cint, (Cstring, Ref{Ptr{Cvoid}}), fcidump_in, fd))
  orbs = Cint.(collect(freeze_orbs))
  st = isempty(orbs) ? 0 : ccall((:ecg_fcidump_freeze, LIB), Cint, (Ptr{Cvoid}, Ptr{Cint}, Cint, Ptr{Cint}, Ptr{Cint}),
                                 fd[], orbs, length(orbs), C_NULL, C_NULL)
  if st == 0
    st = ccall((:ecg_fcidump_write, LIB), Cint, (Ptr{Cvoid}, Cstring, Cdouble), fd[], fcidump_out, tol)
  end
  ccall((:ecg_fcidump_free, LIB), Cint, (Ptr{Cvoid},), fd[])
  _check(nothing, st)
  return nothing
end

"""
    df_freeze(h, BL, BR, e0, nocc, freeze_orbs) -> (h, BL, BR, e0, nocc, nvirt)

Frozen orbitals on DF integrals (`ecg_df_freeze`): `freeze_core!` / `freeze_nvirt!` after the
Fock matrix of all occupied orbitals (src/cc/ccdriver.jl:93-98).  `BR === nothing` for
symmetric factors.  The result feeds `ECGContext(nocc, h, BL; BR=BR, e0=e0)`.
"""
function df_freeze(h::Matrix{Float64}, BL::Array{Float64,3}, BR, e0::Real, nocc::Integer, freeze_orbs)
  n, _, naux = size(BL)
  orbs = Cint.(collect(freeze_orbs))
  nk, nc, e = Ref{Cint}(0), Ref{Cint}(0), Ref{Cdouble}(0.0)
  brp = BR === nothing ? Ptr{Cdouble}(C_NULL) : pointer(BR)
  st = GC.@preserve h BL BR orbs ccall((:ecg_df_freeze, LIB), Cint, (Cint, Cint, Cint, Ptr{Cdouble}, Ptr{Cdouble},
           Ptr{Cdouble}, Cdouble, Ptr{Cint}, Cint, Ref{Cint}, Ref{Cint}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble},
           Ref{Cdouble}), n, nocc, naux, h, BL, brp, e0, orbs, length(orbs), nk, nc,
           C_NULL, C_NULL, C_NULL, e)
  _check(nothing, st)
  m = nk[]
  ho = zeros(m, m); blo = zeros(m, m, naux); bro = BR === nothing ? nothing : zeros(m, m, naux)
  st = GC.@preserve h BL BR orbs ho blo bro ccall((:ecg_df_freeze, LIB), Cint, (Cint, Cint, Cint, Ptr{Cdouble},
           Ptr{Cdouble}, Ptr{Cdouble}, Cdouble, Ptr{Cint}, Cint, Ref{Cint}, Ref{Cint}, Ptr{Cdouble}, Ptr{Cdouble},
           Ptr{Cdouble}, Ref{Cdouble}), n, nocc, naux, h, BL, brp, e0, orbs, length(orbs), nk, nc,
           ho, blo, bro === nothing ? Ptr{Cdouble}(C_NULL) : pointer(bro), e)
  _check(nothing, st)
  no = nocc - nc[]
  return ho, blo, bro, e[], no, m - no
end

end # module
