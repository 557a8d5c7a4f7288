# QED2dHIP.jl -- the reference module's hot-path entry points on libqed2d_hip
# (MI355X / gfx950) through ccall.
#
# A maintainer adds this file next to src/QED2d.jl and replaces
#     using QED2d
# by
#     include("QED2dHIP.jl"); using .QED2dHIP
# in the drivers (main/*.jl) and tests.  QED2dHIP re-exports the reference's
# types (LattParm, KernelParm, RHMCParm, Integrators, Leapfrog, OMF4), so the
# names a driver uses come from one module.  Every function below keeps the
# reference's name, argument order and return value (src/QED2d.jl:27-54); the
# reference file:line it replaces is cited at each definition.
#
# Fields: the library takes DEVICE pointers to the reference's own column-major
# arrays ([L1, L2, 2] ComplexF64 spinors and links, Float64 momenta and forces),
# so no copy or layout change happens at the boundary.  On an MI355X the arrays
# are AMDGPU.jl ROCArrays (CUDA.jl does not run there); `dptr` also accepts a
# CuArray's CuPtr, so code that still builds CuArrays compiles, but the library
# only runs on an AMD GPU.
#
# Kernels: the reference launches seven functions as GPU kernels,
#     CUDA.@cuda threads=kprm.threads blocks=kprm.blocks gamm5Dw(so, U, si, am0, prm)
# (gamm5Dw, gamm5, tr_dQwdU, krnl_force!, updt_fld!, plaquette!, qtop!).  A host
# function cannot be launched by CUDA.@cuda, so those call sites drop the
# launch macro and call the host function of the same name and arguments:
#     gamm5Dw(so, U, si, am0, prm)
# The call sites are src/QED2dHMC.jl:76-78,435,444,463-475 (replaced wholesale
# when HMC!, leapfrog! and update_momenta! come from here),
# test/rhmc_pf_forcetest.jl:60-65 and test/rhmc_zolovarev_bound.jl:35-37.
#
# Errors: every ccall returns a status; a nonzero status becomes
# Base.error(qed2d_last_error(ctx)), where the reference throws
# (src/QED2dSolver.jl:40-43,99-102).
#
# Random numbers: the library draws momenta, pseudofermions and the
# Metropolis uniform from counter-based Philox streams keyed by (seed, chain)
# and indexed by the trajectory (CURAND's streams cannot be reproduced off
# CUDA).  `seed!(s; chain)` replaces Random.seed!(CUDA.default_rng(), s) of
# main/hmc-nf2.jl:45; each HMC! call advances the trajectory index.
module QED2dHIP

import QED2d
using QED2d: LattParm, KernelParm, RHMCParm, Integrators, Leapfrog, OMF4

export LattParm, KernelParm, RHMCParm, Integrators, Leapfrog, OMF4
export gamm5Dw, gamm5, Dw, gamm5Dw_sqr, gamm5Dw_sqr_msq, gamm5Dw_sqr_musq, gamm5Dw_sqr_sqr,
       gamm5Dw_sqr_sqr_musq, tr_dQwdU, krnl_force!, updt_fld!, plaquette!, qtop!,
       Action, Plaquette, Qtop, Hamiltonian, CG, MultiCG, R, generate_pseudofermion!,
       power_method, get_rhmc_params, LuscherZ, LuscherZp, reweighting_factor,
       HMC!, OMF4!, leapfrog!, update_momenta!, unfold_fld!, unfold_reflect!,
       save_gauge, load_gauge, seed!

const lib = "libqed2d_hip"        # on LD_LIBRARY_PATH, or an absolute path

# include/qed2d.h constants
const QED2D_OP_SQR_MSQ = Cint(0)
const QED2D_OP_SQR_MUSQ = Cint(1)
const QED2D_PREC_F64 = Cint(0)
const QED2D_PREC_MIXED = Cint(1)

# src/integrators/integrators.jl:1-3 -> QED2D_LEAPFROG / QED2D_OMF4
integrator_id(::Leapfrog) = Cint(0)
integrator_id(::OMF4) = Cint(1)

# Device address of a field: a ROCArray gives a Ptr, a CuArray a CuPtr; both
# are 64-bit addresses.
dptr(x) = Ptr{Cvoid}(reinterpret(UInt, pointer(x)))

# The stream every context runs on (hipStream_t; C_NULL is the default stream).
# With AMDGPU.jl: QED2dHIP.set_stream!(AMDGPU.stream().stream).
const STREAM = Ref{Ptr{Cvoid}}(C_NULL)
set_stream!(s::Ptr{Cvoid}) = (STREAM[] = s; nothing)

# One context per (L1, L2, beta, fermion boundary flags) -- the counterpart of
# LattParm (src/QED2d.jl:5-8); it owns the library's scratch fields.
const CTX = Dict{Tuple{Int,Int,Float64,Int,UInt},Ptr{Cvoid}}()
const BC = Ref{Cint}(0)          # 1: antiperiodic fermion time BC (extension)

function ctx(L1::Integer, L2::Integer, beta::Real)
    key = (Int(L1), Int(L2), Float64(beta), Int(BC[]), UInt(STREAM[]))
    get!(CTX, key) do
        h = Ref{Ptr{Cvoid}}(C_NULL)
        rc = ccall((:qed2d_ctx_create, lib), Cint,
                   (Ptr{Ptr{Cvoid}}, Cint, Cint, Cdouble, Cint, Ptr{Cvoid}),
                   h, L1, L2, beta, BC[], STREAM[])
        rc == 0 || Base.error("qed2d_ctx_create failed ($rc)")
        h[]
    end
end
ctx(prm::LattParm) = ctx(prm.iL[1], prm.iL[2], prm.beta)
ctx_of(U) = ctx(size(U, 1), size(U, 2), 0.0)   # kernels that take no LattParm

function chk(c::Ptr{Cvoid}, rc::Cint)
    rc == 0 && return nothing
    msg = unsafe_string(ccall((:qed2d_last_error, lib), Cstring, (Ptr{Cvoid},), c))
    Base.error(msg)
end

# ---- random streams ----------------------------------------------------------
const RNG = Ref((seed=UInt64(1234), chain=UInt32(0), traj=UInt32(0)))
seed!(s::Integer; chain::Integer=0, traj::Integer=0) =
    (RNG[] = (seed=UInt64(s), chain=UInt32(chain), traj=UInt32(traj)); nothing)
function next_traj!()
    r = RNG[]
    RNG[] = (seed=r.seed, chain=r.chain, traj=r.traj + UInt32(1))
    return r
end

# ---- Dirac operator (src/QED2dDirac.jl) --------------------------------------
# gamm5Dw(so, U, si, am0, prm) kernel (src/QED2dDirac.jl:2-33)
function gamm5Dw(so, U, si, am0, prm::LattParm)
    c = ctx(prm)
    chk(c, ccall((:qed2d_gamm5Dw, lib), Cint,
                 (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cdouble),
                 c, dptr(so), dptr(U), dptr(si), am0))
end

# gamm5(so, si, am0, prm) kernel (src/QED2dDirac.jl:130-146); am0 is unused there too
function gamm5(so, si, am0, prm::LattParm)
    c = ctx(prm)
    chk(c, ccall((:qed2d_gamm5, lib), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                 c, dptr(so), dptr(si)))
end

# Dw(so, U, si, am0, prm, kprm) (src/QED2dDirac.jl:157-169)
function Dw(so, U, si, am0::Float64, prm::LattParm, kprm::KernelParm)
    c = ctx(prm)
    chk(c, ccall((:qed2d_Dw, lib), Cint,
                 (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cdouble),
                 c, dptr(so), dptr(U), dptr(si), am0))
end

# gamm5Dw_sqr(so, U, si, am0, prm, kprm) (src/QED2dDirac.jl:35-46)
function gamm5Dw_sqr(so, U, si, am0::Float64, prm::LattParm, kprm::KernelParm)
    c = ctx(prm)
    chk(c, ccall((:qed2d_gamm5Dw_sqr, lib), Cint,
                 (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cdouble),
                 c, dptr(so), dptr(U), dptr(si), am0))
end

# gamm5Dw_sqr_msq(so, tmp, U, si, am0, prm, kprm) (src/QED2dDirac.jl:69-81)
function gamm5Dw_sqr_msq(so, tmp, U, si, am0::Float64, prm::LattParm, kprm::KernelParm)
    c = ctx(prm)
    chk(c, ccall((:qed2d_gamm5Dw_sqr_msq, lib), Cint,
                 (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cdouble),
                 c, dptr(so), dptr(tmp), dptr(U), dptr(si), am0))
end

# gamm5Dw_sqr_musq(so, tmp, U, si, am0, mu_j, prm, kprm) (src/QED2dDirac.jl:83-95)
function gamm5Dw_sqr_musq(so, tmp, U, si, am0::Float64, mu_j::Float64, prm::LattParm,
                          kprm::KernelParm)
    c = ctx(prm)
    chk(c, ccall((:qed2d_gamm5Dw_sqr_musq, lib), Cint,
                 (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Cdouble),
                 c, dptr(so), dptr(tmp), dptr(U), dptr(si), am0, mu_j))
end

# gamm5Dw_sqr_sqr(so, U, si, am0, prm, kprm) (src/QED2dDirac.jl:48-67)
function gamm5Dw_sqr_sqr(so, U, si, am0::Float64, prm::LattParm, kprm::KernelParm)
    c = ctx(prm)
    chk(c, ccall((:qed2d_gamm5Dw_sqr_sqr, lib), Cint,
                 (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cdouble),
                 c, dptr(so), dptr(U), dptr(si), am0))
end

# gamm5Dw_sqr_sqr_musq(so, tmp, U, si, am0, mu_j, prm, kprm) (src/QED2dDirac.jl:97-117)
function gamm5Dw_sqr_sqr_musq(so, tmp, U, si, am0::Float64, mu_j::Float64, prm::LattParm,
                              kprm::KernelParm)
    c = ctx(prm)
    chk(c, ccall((:qed2d_gamm5Dw_sqr_sqr_musq, lib), Cint,
                 (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Cdouble),
                 c, dptr(so), dptr(tmp), dptr(U), dptr(si), am0, mu_j))
end

# tr_dQwdU(frc, U, X, g5DwX, prm) kernel (src/QED2dDirac.jl:172-199)
function tr_dQwdU(frc, U, X, g5DwX, prm::LattParm)
    c = ctx(prm)
    chk(c, ccall((:qed2d_tr_dQwdU, lib), Cint,
                 (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                 c, dptr(frc), dptr(U), dptr(X), dptr(g5DwX)))
end

# The function the reference passes as `A` (src/QED2dSolver.jl:20,65,79): the
# reference's operator or this module's of the same name.
op_id(::typeof(QED2d.gamm5Dw_sqr_msq)) = QED2D_OP_SQR_MSQ
op_id(::typeof(QED2d.gamm5Dw_sqr_musq)) = QED2D_OP_SQR_MUSQ
op_id(::typeof(gamm5Dw_sqr_msq)) = QED2D_OP_SQR_MSQ
op_id(::typeof(gamm5Dw_sqr_musq)) = QED2D_OP_SQR_MUSQ
op_id(A) = Base.error("QED2dHIP: CG operator $A has no library counterpart; " *
                      "call QED2d.CG for it")

# ---- gauge action, forces, link update (src/QED2dAction.jl) -------------------
# plaquette!(plx, U, prm) kernel (src/QED2dAction.jl:2-16)
function plaquette!(plx, U, prm::LattParm)
    c = ctx(prm)
    chk(c, ccall((:qed2d_plaquette_field, lib), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                 c, dptr(plx), dptr(U)))
end

# qtop!(plx, U, prm) kernel (src/QED2dAction.jl:18-32)
function qtop!(plx, U, prm::LattParm)
    c = ctx(prm)
    chk(c, ccall((:qed2d_qtop_field, lib), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                 c, dptr(plx), dptr(U)))
end

# krnl_force!(frc1, frc2, U, prm) kernel (src/QED2dAction.jl:34-53)
function krnl_force!(frc1, frc2, U, prm::LattParm)
    c = ctx(prm)
    chk(c, ccall((:qed2d_krnl_force, lib), Cint,
                 (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                 c, dptr(frc1), dptr(frc2), dptr(U)))
end

# updt_fld!(U, mom, eps) kernel (src/QED2dAction.jl:55-64): no LattParm there,
# the lattice comes from size(U)
function updt_fld!(U, mom, eps)
    c = ctx_of(U)
    chk(c, ccall((:qed2d_updt_fld, lib), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cdouble),
                 c, dptr(U), dptr(mom), eps))
end

# ---- scalars (src/QED2dHMC.jl:198-237) ------------------------------------------
# Action(U, prm, kprm) (src/QED2dHMC.jl:206-215)
function Action(U, prm::LattParm, kprm::KernelParm)
    c = ctx(prm)
    v = Ref{Cdouble}(0.0)
    chk(c, ccall((:qed2d_action, lib), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cdouble}),
                 c, dptr(U), v))
    return v[]
end
# Plaquette(U, prm, kprm) (src/QED2dHMC.jl:217-226)
function Plaquette(U, prm::LattParm, kprm::KernelParm)
    c = ctx(prm)
    v = Ref{Cdouble}(0.0)
    chk(c, ccall((:qed2d_plaquette, lib), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cdouble}),
                 c, dptr(U), v))
    return v[]
end
# Qtop(U, prm, kprm) (src/QED2dHMC.jl:228-237)
function Qtop(U, prm::LattParm, kprm::KernelParm)
    c = ctx(prm)
    v = Ref{Cdouble}(0.0)
    chk(c, ccall((:qed2d_qtop, lib), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cdouble}),
                 c, dptr(U), v))
    return v[]
end
# Hamiltonian(mom, U, prm, kprm) (src/QED2dHMC.jl:198-203)
function Hamiltonian(mom, U, prm::LattParm, kprm::KernelParm)
    c = ctx(prm)
    h = Ref{Cdouble}(0.0)
    chk(c, ccall((:qed2d_hamiltonian, lib), Cint,
                 (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cdouble}), c, dptr(mom), dptr(U), h))
    return h[]
end

# ---- solvers (src/QED2dSolver.jl, src/QED2dRHMC.jl) -------------------------------
# CG(so, U, si, am0, mu_mass, maxiter, eps, A, prm, kprm; so_as_guess)
# (src/QED2dSolver.jl:49-105); `precision` selects FP32 iterations with FP64
# reliable updates (QED2D_PREC_MIXED), stopping on the same FP64 residual.
function CG(so, U, si, am0, mu_mass, maxiter, eps, A, prm::LattParm, kprm::KernelParm;
            so_as_guess::Bool=false, precision::Integer=QED2D_PREC_F64)
    c = ctx(prm)
    it = Ref{Cint}(0)
    chk(c, ccall((:qed2d_cg, lib), Cint,
                 (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Cdouble, Cint, Cint,
                  Cdouble, Cint, Cint, Ptr{Cint}),
                 c, dptr(so), dptr(U), dptr(si), am0, mu_mass, op_id(A), maxiter, eps,
                 so_as_guess, precision, it))
    return Int(it[])
end
# CG(so, U, si, am0, maxiter, eps, A, prm, kprm) (src/QED2dSolver.jl:3-45)
CG(so, U, si, am0, maxiter::Integer, eps, A, prm::LattParm, kprm::KernelParm;
   precision::Integer=QED2D_PREC_F64) =
    CG(so, U, si, am0, 0.0, maxiter, eps, A, prm, kprm; precision=precision)

# MultiCG(so, U, si, am0, maxiter, eps, A, rprm, prm, kprm) (src/QED2dRHMC.jl:281-290);
# A is gamm5Dw_sqr_musq there, the shifted operator the library solves
function MultiCG(so, U, si, am0, maxiter, eps, A, rprm::RHMCParm, prm::LattParm,
                 kprm::KernelParm)
    op_id(A) == QED2D_OP_SQR_MUSQ || Base.error("MultiCG: A must be gamm5Dw_sqr_musq")
    c = ctx(prm)
    chk(c, ccall((:qed2d_multicg, lib), Cint,
                 (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Cint, Ptr{Cdouble},
                  Ptr{Cdouble}, Cint, Cdouble),
                 c, dptr(so), dptr(U), dptr(si), am0, rprm.n, rprm.mu, rprm.rho, maxiter, eps))
end

# R(so, U, si, am0, maxiter, eps, A, rprm, prm, kprm) (src/QED2dRHMC.jl:261-271)
function R(so, U, si, am0, maxiter, eps, A, rprm::RHMCParm, prm::LattParm, kprm::KernelParm)
    op_id(A) == QED2D_OP_SQR_MUSQ || Base.error("R: A must be gamm5Dw_sqr_musq")
    c = ctx(prm)
    chk(c, ccall((:qed2d_R, lib), Cint,
                 (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Cint, Ptr{Cdouble},
                  Ptr{Cdouble}, Cdouble, Cdouble, Cint, Cdouble),
                 c, dptr(so), dptr(U), dptr(si), am0, rprm.n, rprm.mu, rprm.rho, rprm.A,
                 rprm.r_b, maxiter, eps))
end

# generate_pseudofermion!(F, U, X, am0, CGmaxiter, CGtol, prm, kprm, rprm)
# (src/QED2dRHMC.jl:298-319)
function generate_pseudofermion!(F, U, X, am0, CGmaxiter, CGtol, prm, kprm, rprm::RHMCParm)
    c = ctx(prm)
    chk(c, ccall((:qed2d_generate_pseudofermion, lib), Cint,
                 (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Cint, Ptr{Cdouble},
                  Ptr{Cdouble}, Cint, Cdouble),
                 c, dptr(F), dptr(U), dptr(X), am0, rprm.n, rprm.mu, rprm.nu, CGmaxiter,
                 CGtol))
end

# power_method(U, am0, prm, kprm; iter) (src/QED2dRHMC.jl:135-170).  The start
# vector (CUDA.randn at :137) is the pseudofermion (flavour 15) of the next
# trajectory index, as the Python mirror draws it; the results are returned as
# ComplexF64 like the reference's dot-product ratios.
function power_method(U, am0, prm, kprm; iter::Int64=1000)
    c = ctx(prm)
    b0 = similar(U)
    r = next_traj!()
    chk(c, ccall((:qed2d_fill_pseudofermion, lib), Cint,
                 (Ptr{Cvoid}, Ptr{Cvoid}, UInt64, UInt32, UInt32, UInt32),
                 c, dptr(b0), r.seed, r.chain, r.traj, UInt32(15)))
    lmin, lmax = Ref{Cdouble}(0.0), Ref{Cdouble}(0.0)
    chk(c, ccall((:qed2d_power_method, lib), Cint,
                 (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Cint, Ptr{Cdouble}, Ptr{Cdouble}),
                 c, dptr(U), dptr(b0), am0, iter, lmin, lmax))
    return ComplexF64(lmin[]), ComplexF64(lmax[])
end

# get_rhmc_params(n, r_a, r_b; reweighting_N, reweighting_Taylor)
# (src/QED2dRHMC.jl:93-110): Zolotarev coefficients on the host
function get_rhmc_params(n_rhmc::Int64, r_a_rhmc, r_b_rhmc; reweighting_N::Int64=1,
                         reweighting_Taylor::Int64=5)
    rho, mu, nu = zeros(n_rhmc), zeros(n_rhmc), zeros(n_rhmc)
    e, A, d = Ref{Cdouble}(0.0), Ref{Cdouble}(0.0), Ref{Cdouble}(0.0)
    r_b = Float64(real(r_b_rhmc))
    rc = ccall((:qed2d_zolotarev, lib), Cint,
               (Cint, Cdouble, Cdouble, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble},
                Ptr{Cdouble}, Ptr{Cdouble}),
               n_rhmc, Float64(real(r_a_rhmc)), r_b, e, A, d, rho, mu, nu)
    rc == 0 || Base.error("get_rhmc_params: bad arguments (n = $n_rhmc)")
    return RHMCParm(r_b, n_rhmc, e[], A[], rho, mu, nu, d[], reweighting_N, reweighting_Taylor)
end
# (src/QED2dRHMC.jl:112-121)
get_rhmc_params(n_rhmc::Array, r_a_rhmc::Array, r_b_rhmc::Array; reweighting_N::Int64=1,
                reweighting_Taylor::Int64=4) =
    RHMCParm[get_rhmc_params(n_rhmc[j], r_a_rhmc[j], r_b_rhmc[j]) for j in 1:length(n_rhmc)]

# LuscherZp(Z, p, U, X_in, am0, CGmaxiter, CGtol, rprm, prm, kprm) (src/QED2dRHMC.jl:189-201)
function LuscherZp(Z, p::Int64, U, X_in, am0, CGmaxiter, CGtol, rprm::RHMCParm, prm::LattParm,
                   kprm::KernelParm)
    c = ctx(prm)
    chk(c, ccall((:qed2d_luscher_z, lib), Cint,
                 (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Cint,
                  Ptr{Cdouble}, Ptr{Cdouble}, Cdouble, Cdouble, Cint, Cdouble),
                 c, dptr(Z), p, dptr(U), dptr(X_in), am0, rprm.n, rprm.mu, rprm.rho, rprm.A,
                 rprm.r_b, CGmaxiter, CGtol))
end
# LuscherZ(Z, U, X_in, am0, CGmaxiter, CGtol, rprm, prm, kprm) (src/QED2dRHMC.jl:174-186)
LuscherZ(Z, U, X_in, am0, CGmaxiter, CGtol, rprm::RHMCParm, prm::LattParm, kprm::KernelParm) =
    LuscherZp(Z, 1, U, X_in, am0, CGmaxiter, CGtol, rprm, prm, kprm)

# reweighting_factor(U, am0, CGmaxiter, CGtol, prm, kprm, rprm) (src/QED2dRHMC.jl:210-239);
# the Gaussian fields (CUDA.randn at :218) come from the next trajectory index's stream
function reweighting_factor(U, am0, CGmaxiter, CGtol, prm::LattParm, kprm::KernelParm,
                            rprm::RHMCParm)
    c = ctx(prm)
    W = Ref{Cdouble}(0.0)
    r = next_traj!()
    chk(c, ccall((:qed2d_reweighting_factor, lib), Cint,
                 (Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Cint, Ptr{Cdouble}, Ptr{Cdouble}, Cdouble,
                  Cdouble, Cint, Cint, Cint, Cdouble, UInt64, UInt32, UInt32, Ptr{Cdouble}),
                 c, dptr(U), am0, rprm.n, rprm.mu, rprm.rho, rprm.A, rprm.r_b,
                 rprm.reweighting_N, rprm.reweighting_Taylor, CGmaxiter, CGtol, r.seed,
                 r.chain, r.traj, W))
    return W[]
end
# the product over fermions (src/QED2dRHMC.jl:242-251)
reweighting_factor(U, am0::Array{Float64}, CGmaxiter, CGtol, prm::LattParm, kprm::KernelParm,
                   rprm::Array{RHMCParm}) =
    prod(reweighting_factor(U, am0[j], CGmaxiter, CGtol, prm, kprm, rprm[j])
         for j in 1:length(am0))

# ---- molecular dynamics (src/QED2dHMC.jl) -------------------------------------------
# OMF4!(mom, U, eps, ns, prm, kprm), pure gauge (src/QED2dHMC.jl:239-305)
function OMF4!(mom, U, eps, ns, prm::LattParm, kprm::KernelParm)
    c = ctx(prm)
    chk(c, ccall((:qed2d_omf4_pure, lib), Cint,
                 (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Cint),
                 c, dptr(mom), dptr(U), eps, ns))
end

function md_nf2(itg::Integrators, mom, U, X, F, g5DX, am0, eps, ns, maxiter, tol,
                prm::LattParm; precision::Integer=QED2D_PREC_F64)
    c = ctx(prm)
    it = Ref{Cint}(0)
    chk(c, ccall((:qed2d_md_nf2, lib), Cint,
                 (Ptr{Cvoid}, Cint, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid},
                  Cdouble, Cdouble, Cint, Cint, Cdouble, Cint, Ptr{Cint}),
                 c, integrator_id(itg), dptr(mom), dptr(U), dptr(X), dptr(F), dptr(g5DX),
                 am0, eps, ns, maxiter, tol, precision, it))
    return nothing
end
# OMF4!(mom, U, X, F, g5DX, am0, eps, ns, maxiter, tol, prm, kprm), Nf=2
# (src/QED2dHMC.jl:307-422)
OMF4!(mom, U, X, F, g5DX, am0, eps, ns, maxiter, tol, prm::LattParm, kprm::KernelParm;
      precision::Integer=QED2D_PREC_F64) =
    md_nf2(OMF4(), mom, U, X, F, g5DX, am0, eps, ns, maxiter, tol, prm; precision=precision)
# leapfrog!(mom, U, X, F, g5DX, am0, eps, ns, CGmaxiter, tol, prm, kprm), Nf=2
# (src/QED2dHMC.jl:426-449)
leapfrog!(mom, U, X, F, g5DX, am0, eps, ns, CGmaxiter, tol, prm::LattParm, kprm::KernelParm;
          precision::Integer=QED2D_PREC_F64) =
    md_nf2(Leapfrog(), mom, U, X, F, g5DX, am0, eps, ns, CGmaxiter, tol, prm;
           precision=precision)

# update_momenta!(mom, U, X, F, g5DX, am0, eps, ns, maxiter, tol, prm, kprm), Nf=2
# (src/QED2dHMC.jl:452-481): X = (Q^2)^-1 F, g5DX = Q X, mom += eps (F_gauge + F_f)
function update_momenta!(mom, U, X, F, g5DX, am0, eps, ns, maxiter, tol, prm::LattParm,
                         kprm::KernelParm; precision::Integer=QED2D_PREC_F64)
    CG(X, U, F, am0, maxiter, tol, gamm5Dw_sqr_msq, prm, kprm; precision=precision)
    gamm5Dw(g5DX, U, X, am0, prm)
    c = ctx(prm)
    chk(c, ccall((:qed2d_force_mom_update, lib), Cint,
                 (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cdouble),
                 c, dptr(mom), dptr(U), dptr(X), dptr(g5DX), eps))
end

# ---- HMC trajectories (src/QED2dHMC.jl:19-195) ----------------------------------------
function report(acc, a, hi, hf)
    push!(acc, Int(a))
    pacc = count(acc .== 1) / length(acc)
    if a == 0
        @info("    REJECT: Energy [inital: $hi; final: $hf; difference: $(hf - hi)]; Pacc = $pacc")
    else
        @info("    ACCEPT: Energy [inital: $hi; final: $hf; difference: $(hf - hi)]; Pacc = $pacc")
    end
    return nothing
end

# HMC!(U, eps, ns, acc, prm, kprm; integrator, qzero), pure gauge (src/QED2dHMC.jl:19-60)
function HMC!(U, eps, ns, acc, prm::LattParm, kprm::KernelParm;
              integrator::Integrators=OMF4(), qzero=false)
    c = ctx(prm)
    r = next_traj!()
    hi, hf, a = Ref{Cdouble}(0.0), Ref{Cdouble}(0.0), Ref{Cint}(0)
    chk(c, ccall((:qed2d_hmc_pure, lib), Cint,
                 (Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Cint, Cint, Cint, UInt64, UInt32, UInt32,
                  Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cint}),
                 c, dptr(U), eps, ns, integrator_id(integrator), qzero, r.seed, r.chain,
                 r.traj, hi, hf, a))
    report(acc, a[], hi[], hf[])
end

# HMC!(U, am0, eps, ns, acc, CGmaxiter, tol, prm, kprm; integrator, qzero), Nf=2
# (src/QED2dHMC.jl:64-113).  precision / acc_tol: the MD-force solver and the
# acceptance re-solve (include/qed2d.h); the defaults are the reference's.
function HMC!(U, am0::Real, eps, ns, acc, CGmaxiter, tol, prm::LattParm, kprm::KernelParm;
              integrator::Integrators=Leapfrog(), qzero=false,
              precision::Integer=QED2D_PREC_F64, acc_tol::Real=0.0)
    c = ctx(prm)
    r = next_traj!()
    hi, hf, a, it = Ref{Cdouble}(0.0), Ref{Cdouble}(0.0), Ref{Cint}(0), Ref{Cint}(0)
    chk(c, ccall((:qed2d_hmc_nf2, lib), Cint,
                 (Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Cdouble, Cint, Cint, Cint, Cdouble, Cint,
                  Cint, Cdouble, UInt64, UInt32, UInt32, Ptr{Cdouble}, Ptr{Cdouble},
                  Ptr{Cint}, Ptr{Cint}),
                 c, dptr(U), am0, eps, ns, integrator_id(integrator), CGmaxiter, tol, qzero,
                 precision, acc_tol, r.seed, r.chain, r.traj, hi, hf, a, it))
    report(acc, a[], hi[], hf[])
end

# HMC!(U, am0::Array, eps, ns, acc, CGmaxiter, tol, prm, kprm, rprm::Array{RHMCParm};
#      integrator, qzero, reversibility_test), RHMC (src/QED2dHMC.jl:116-195).
# OMF4() is an extension here: the reference's MD! has no RHMC OMF4 method.
function HMC!(U, am0::Array, eps, ns, acc, CGmaxiter, tol, prm::LattParm, kprm::KernelParm,
              rprm::Array{RHMCParm}; integrator::Integrators=Leapfrog(), qzero=false,
              reversibility_test=false)
    c = ctx(prm)
    r = next_traj!()
    nf = length(am0)
    m0 = Float64.(am0)
    npole = Cint[p.n for p in rprm]
    mus = [pointer(p.mu) for p in rprm]
    nus = [pointer(p.nu) for p in rprm]
    rhos = [pointer(p.rho) for p in rprm]
    hi, hf, a = Ref{Cdouble}(0.0), Ref{Cdouble}(0.0), Ref{Cint}(0)
    GC.@preserve rprm begin
        chk(c, ccall((:qed2d_hmc_rhmc, lib), Cint,
                     (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cdouble}, Cint, Ptr{Cint}, Ptr{Ptr{Cdouble}},
                      Ptr{Ptr{Cdouble}}, Ptr{Ptr{Cdouble}}, Cdouble, Cint, Cint, Cint, Cdouble,
                      Cint, Cint, UInt64, UInt32, UInt32, Ptr{Cdouble}, Ptr{Cdouble},
                      Ptr{Cint}),
                     c, dptr(U), m0, nf, npole, mus, nus, rhos, eps, ns,
                     integrator_id(integrator), CGmaxiter, tol, qzero, reversibility_test,
                     r.seed, r.chain, r.traj, hi, hf, a))
    end
    reversibility_test && return nothing        # (src/QED2dHMC.jl:159-164)
    report(acc, a[], hi[], hf[])
end

# ---- fields and files --------------------------------------------------------------------
# unfold_fld!(Uin, prm, kprm) / unfold_reflect!(Uin, prm, kprm) (src/QED2dFields.jl:1-50)
function unfold(Uin, prm::LattParm, reflect::Integer)
    c = ctx(prm)
    Uout = similar(Uin, 2 * prm.iL[1], 2 * prm.iL[2], 2)
    chk(c, ccall((:qed2d_unfold, lib), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint),
                 c, dptr(Uout), dptr(Uin), reflect))
    return Uout
end
unfold_fld!(Uin, prm::LattParm, kprm::KernelParm) = unfold(Uin, prm, 0)
unfold_reflect!(Uin, prm::LattParm, kprm::KernelParm) = unfold(Uin, prm, 1)

# save_gauge(U, file, prm) / load_gauge(U, file, prm) (src/QED2dLoadSave.jl:1-19):
# the same text format (writedlm of hcat(real, imag), column-major)
function save_gauge(U, file::String, prm::LattParm)
    c = ctx(prm)
    chk(c, ccall((:qed2d_save_gauge, lib), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Cstring),
                 c, dptr(U), file))
end
function load_gauge(U, file::String, prm::LattParm)
    c = ctx(prm)
    chk(c, ccall((:qed2d_load_gauge, lib), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Cstring),
                 c, dptr(U), file))
end

end # module
