"""
    ElemCoAMD

Julia binding of `libelemco_hip.so`, the MI355X (gfx950) library of the
closed-shell coupled-cluster path of ElemCo.jl.  Every wrapper is a `ccall` of
one entry point of `include/elemco_hip.h`; arrays are passed as Julia
column-major `Float64` arrays (the layout the library expects), and a negative
status is raised with `error(...)`, the reference's convention (e.g.
`src/cc/cc.jl:1364`).

The drop-in methods at the end keep the reference's signatures:

- `calc_cc_resid(ctx, T1, T2; dc)`            ↔ `src/cc/cc.jl:1343-1489`
- `calc_ccsd_vector_times_Jacobian(ctx, …)`  ↔ `src/cc/cc_lagrange.jl:9-212`
- `calc_1RDM(ctx, U1, U2, T1, T2)`          ↔ `src/cc/cc_lagrange.jl:793-837`
- `calc_pertT(ctx, T1, T2; U1, U2)`          ↔ `src/cc/cc_triples.jl:39-288`
- `calc_cc(ctx, method)`                     ↔ `src/cc/cc.jl:2451-2516`
- `calc_lm_cc(ctx, T1, T2, method)`          ↔ `src/cc/cc_lagrange.jl:1120-1223`

`ctx` stands for the reference's `EC::ECInfo` on this path: it owns the
integrals resident on one GPU.  Build it from MO density-fitting factors
`B[p,q,Q]` (the layout of `mmL`, `src/integrals/dftools.jl:102`) and the
one-electron integrals, or straight from an FCIDUMP file (`ECGContext(path)`).
"""
module ElemCoAMD

export ECGContext, write_fcidump, df_freeze, calc_cc_resid, calc_ccsd_vector_times_Jacobian, calc_1RDM, calc_pertT, calc_cc, calc_lm_cc,
       energy_grad, SolverOpts, default_opts, ECG_FP64, ECG_FP32, ECG_BF16X3

const LIB = get(ENV, "ELEMCO_HIP_LIB", joinpath(@__DIR__, "..", "libelemco_hip.so"))

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
end

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

_z1(ctx) = zeros(ctx.nvirt, ctx.nocc)
_z2(ctx) = zeros(ctx.nvirt, ctx.nvirt, ctx.nocc, ctx.nocc)

"`calc_cc_resid(EC, T1, T2; dc)` (src/cc/cc.jl:1343-1489): returns freshly allocated `(R1, R2)`."
function calc_cc_resid(ctx::ECGContext, T1::Matrix{Float64}, T2::Array{Float64,4}; dc=false, precision=ECG_FP64)
  R1, R2 = _z1(ctx), _z2(ctx)
  st = ccall((:ecg_cc_residual, LIB), Cint,
             (Ptr{Cvoid}, Cint, Cint, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}),
             ctx.handle, dc ? ECG_DCSD : ECG_CCSD, precision, T1, T2, R1, R2)
  _check(ctx, st)
  return R1, R2
end

"`calc_ccsd_vector_times_Jacobian(EC, U1, U2; dc)` (src/cc/cc_lagrange.jl:9-212) at amplitudes T1, T2."
function calc_ccsd_vector_times_Jacobian(ctx::ECGContext, U1, U2, T1, T2; dc=false, precision=ECG_FP64)
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
`calc_pertT` (src/cc/cc_triples.jl:39-288): `Dict("ET3" => E(T), "ET3b" => E[T])`;
Λ-(T) when `U1`, `U2` are given.  `first`/`count` restrict the evaluation to a
range of triplets of the reference's loop order (partial sums add up).
"""
function calc_pertT(ctx::ECGContext, T1, T2; U1=nothing, U2=nothing, precision=ECG_FP64, first=0, count=-1)
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
  return Dict("ET3" => et[], "ET3b" => etb[])
end

"""
`calc_cc(EC, method)` (src/cc/cc.jl:2451-2516) for "ccsd" / "dcsd": iterate on the
device from the MP2 guess; returns `(energies, T1, T2)` with
`energies = Dict("E" => E_H, "ESS", "EOS", "Eproj", "dE", "NormR")` of the last iteration.
"""
function calc_cc(ctx::ECGContext, method::AbstractString; opts=default_opts(), precision=ECG_FP64)
  m = lowercase(method) == "dcsd" ? ECG_DCSD : ECG_CCSD
  T1, T2 = _z1(ctx), _z2(ctx)
  e = zeros(6)
  it, conv = Ref{Cint}(0), Ref{Cint}(0)
  st = ccall((:ecg_cc_solve, LIB), Cint,
             (Ptr{Cvoid}, Cint, Cint, Ref{SolverOpts}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ref{Cint}, Ref{Cint}),
             ctx.handle, m, precision, opts, T1, T2, e, it, conv)
  _check(ctx, st)
  conv[] == 0 && @warn "CC iterations did not converge in $(it[]) iterations"
  return Dict("E" => e[1], "ESS" => e[2], "EOS" => e[3], "Eproj" => e[4], "dE" => e[5], "NormR" => e[6]), T1, T2
end

"`calc_lm_cc(EC, method)` (src/cc/cc_lagrange.jl:1120-1223): Λ multipliers `(U1, U2)` at amplitudes T1, T2."
function calc_lm_cc(ctx::ECGContext, T1, T2, method::AbstractString; opts=default_opts(), precision=ECG_FP64)
  m = lowercase(method) == "dcsd" ? ECG_DCSD : ECG_CCSD
  U1, U2 = _z1(ctx), _z2(ctx)
  it, conv = Ref{Cint}(0), Ref{Cint}(0)
  st = ccall((:ecg_lambda_solve, LIB), Cint,
             (Ptr{Cvoid}, Cint, Cint, Ref{SolverOpts}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble},
              Ref{Cint}, Ref{Cint}),
             ctx.handle, m, precision, opts, T1, T2, U1, U2, it, conv)
  _check(ctx, st)
  conv[] == 0 && @warn "Lambda iterations did not converge in $(it[]) iterations"
  return U1, U2
end

"""
`energy_grad(ctx, T1, T2, U1, U2, method)`: `(dE/dh [n,n], dE/dB [n,n,naux])` of the
correlation energy at converged amplitudes and Λ multipliers (no reference
counterpart; the backward a ChainRulesCore `rrule` of the CC energy would call).
"""
function energy_grad(ctx::ECGContext, T1, T2, U1, U2, method::AbstractString)
  m = lowercase(method) == "dcsd" ? ECG_DCSD : ECG_CCSD
  n = ctx.nocc + ctx.nvirt
  gh, gb = zeros(n, n), zeros(n, n, ctx.naux)
  st = ccall((:ecg_energy_grad, LIB), Cint,
             (Ptr{Cvoid}, Cint, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}),
             ctx.handle, m, T1, T2, U1, U2, gh, gb)
  _check(ctx, st)
  return gh, gb
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
