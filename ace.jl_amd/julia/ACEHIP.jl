"""
    ACEHIP

The evaluator slot of ACE.jl's `LinearACEModel` (src/linearmodel.jl:36-56)
backed by libacehip.so, the MI355X (gfx950) basis-evaluation library of this
repository.  `HIPEvaluator(basis, c)` takes the place of
`ProductEvaluator(basis, c)` (src/evaluator.jl:9-31): the library compiles
its device programs from the basis's own tables (`ace_ctx_create_jl`:
Product1pBasis indices, PIBasisSpec, A2Bmap, the radial OrthPolyBasis) and
evaluates site energies, their gradients, basis values and the
`adjoint_EVAL_D` contraction on the GPU.

Every method ACE.jl dispatches on the evaluator is implemented:
`evaluate`, `grad_config`, `_rrule_evaluate`, `adjoint_EVAL_D`,
`set_params!`, `write_dict` / `read_dict`; and `grad_params` for models that
use this evaluator.  ACE.jl's own `rrule`s for `evaluate` and for its
pullback (src/linearmodel.jl:178-229) call these methods, so first and mixed
second derivatives work unchanged.  `evaluate_batch` / `grad_config_batch`
evaluate many configurations in one call, which is how the GPU should be
used.

Usage:

    using ACE, ACEHIP
    basis = ACE.SymmetricBasis(ACE.Invariant(), B1p, Bsel)
    model = ACEHIP.hip_model(basis, c)           # LinearACEModel(basis, c, HIPEvaluator(basis, c))
    E  = ACE.evaluate(model, cfg)                # Invariant{Float64}
    dE = ACE.grad_config(model, cfg)             # Vector{DState}

The library is found at `ENV["LIBACEHIP"]`, else next to this file
(`../acehip/libacehip.so`).  Arrays cross the C ABI as host memory (the
`ace_*_host` entry points stage them to the device).
"""
module ACEHIP

using ACE
using ACE: LinearACEModel, SymmetricBasis, DState
using SparseArrays
using StaticArrays

export HIPEvaluator, hip_model, evaluate_batch, grad_config_batch

const libacehip = get(ENV, "LIBACEHIP", normpath(joinpath(@__DIR__, "..", "acehip", "libacehip.so")))

const ABI_VERSION = Int32(3)
const PROP_INVARIANT = Int32(0)
const PROP_EUCLIDEAN_VECTOR = Int32(1)
const COMP_CATEGORICAL = Int32(0)
const COMP_RN = Int32(1)
const COMP_YLM = Int32(2)

# `ace_jl_basis` of include/acehip.h, field for field (C layout)
struct AceJlBasis
    abi_version::Int32
    prop::Int32
    nr::Int64
    A::Ptr{Float64}
    B::Ptr{Float64}
    C::Ptr{Float64}
    pl::Int64
    pr::Int64
    tl::Float64
    tr::Float64
    tp::Int64
    r0::Float64
    maxL::Int64
    ncategories::Int64
    ncomp::Int64
    comp_kind::Ptr{Int32}
    nA::Int64
    indices::Ptr{Int64}
    nAA::Int64
    maxorder::Int64
    orders::Ptr{Int64}
    iAA2iA::Ptr{Int64}
    nB::Int64
    colptr::Ptr{Int64}
    rowval::Ptr{Int64}
    nzval::Ptr{Float64}
end

_last_error() = unsafe_string(ccall((:ace_last_error, libacehip), Cstring, ()))
_check(rc) = (rc == 0 || error("libacehip: " * _last_error()); nothing)

# ---------------------------------------------------------------------------
# reading the basis

# The transform of the radial basis: polytransform(p, r0) is the lambda
# "r -> ((1+r0)/(1+r))^p" (src/transforms/distancetransforms.jl:16).
const _POLYTRANSFORM = r"^\s*r\s*->\s*\(\(1\s*\+\s*([-+0-9.eE]+)\)\s*/\s*\(1\s*\+\s*r\)\)\s*\^\s*([0-9]+)\s*$"

function _radial(Rn)
    F = Rn.basis.F                       # chain(λ("rr -> norm(rr)"), transform, OrthPolyBasis)
    J = F[end]
    J isa ACE.OrthPolys.OrthPolyBasis || error("ACEHIP: the radial component must end in an OrthPolyBasis")
    m = match(_POLYTRANSFORM, F[end-1].exstr)
    m === nothing && error("ACEHIP: the radial transform must be polytransform(p, r0), got $(F[end-1].exstr)")
    return J, parse(Int, m.captures[2]), parse(Float64, m.captures[1])
end

function _kind(B)
    B isa ACE.Categorical1pBasis && return COMP_CATEGORICAL
    if B isa ACE.B1pComponent
        B.basis isa ACE.SphericalHarmonics.SHBasis && return COMP_YLM
        if B.basis isa ACE.SChain && B.basis.F[end] isa ACE.OrthPolys.OrthPolyBasis
            return COMP_RN
        end
    end
    error("ACEHIP: unsupported 1p component $(typeof(B)); Rn, Ylm and Categorical1pBasis are supported")
end

# The arrays behind an AceJlBasis (kept alive while the library reads them).
function _tables(basis::SymmetricBasis)
    basis.real === Base.real || error("ACEHIP: only real-valued symmetric bases (isreal = true) are supported")
    T = eltype(basis.A2Bmap)
    prop = T <: ACE.Invariant ? PROP_INVARIANT :
           T <: ACE.EuclideanVector ? PROP_EUCLIDEAN_VECTOR :
           error("ACEHIP: property $(T) is not supported (Invariant, EuclideanVector)")
    B1p = basis.pibasis.basis1p
    kinds = Int32[_kind(B) for B in B1p.bases]
    iRn = findfirst(==(COMP_RN), kinds)
    iY = findfirst(==(COMP_YLM), kinds)
    iC = findfirst(==(COMP_CATEGORICAL), kinds)
    J, tp, r0 = _radial(B1p.bases[iRn])
    maxL = B1p.bases[iY].basis.alp.L
    cat = iC === nothing ? nothing : B1p.bases[iC]
    ncat = cat === nothing ? 0 : length(cat.categories.list)
    indices = collect(reinterpret(Int, B1p.indices))        # [nA][ncomp]
    spec = basis.pibasis.spec
    A2B = basis.A2Bmap
    nzval = collect(reinterpret(Float64, A2B.nzval))         # 1 or 6 Float64 per entry
    arrays = (Vector{Float64}(J.A), Vector{Float64}(J.B), Vector{Float64}(J.C), kinds, indices,
              Vector{Int}(spec.orders), Matrix{Int}(spec.iAA2iA), Vector{Int}(A2B.colptr),
              Vector{Int}(A2B.rowval), nzval)
    return arrays, (prop = prop, J = J, tp = tp, r0 = r0, maxL = maxL, ncat = ncat, cat = cat,
                    nB = size(A2B, 1))
end

function _jl_basis(arrays, info)
    A, B, C, kinds, indices, orders, iAA2iA, colptr, rowval, nzval = arrays
    J = info.J
    return AceJlBasis(ABI_VERSION, info.prop, length(A), pointer(A), pointer(B), pointer(C),
                      J.pl, J.pr, J.tl, J.tr, info.tp, info.r0, info.maxL, info.ncat,
                      length(kinds), pointer(kinds), length(indices) ÷ length(kinds), pointer(indices),
                      length(orders), size(iAA2iA, 2), pointer(orders), pointer(iAA2iA),
                      info.nB, pointer(colptr), pointer(rowval), pointer(nzval))
end

# ---------------------------------------------------------------------------
# the evaluator

"""
    HIPEvaluator(basis::SymmetricBasis, c; device = 0)

A libacehip.so context for `basis` with coefficients `c` on HIP device
`device`: the GPU counterpart of `ProductEvaluator(basis, c)`.
"""
mutable struct HIPEvaluator
    ctx::Ptr{Cvoid}
    prop::Int32
    nB::Int
    categorical::Any      # the Categorical1pBasis (neighbour species), or nothing
    device::Int
end

function HIPEvaluator(basis::SymmetricBasis, c::AbstractVector; device::Integer = 0)
    arrays, info = _tables(basis)
    cv = Vector{Float64}(c)
    length(cv) == info.nB || error("ACEHIP: length(c) != length(basis)")
    out = Ref{Ptr{Cvoid}}(C_NULL)
    a1, a2, a3, a4, a5, a6, a7, a8, a9, a10 = arrays
    GC.@preserve a1 a2 a3 a4 a5 a6 a7 a8 a9 a10 cv begin
        jb = Ref(_jl_basis(arrays, info))
        _check(ccall((:ace_ctx_create_jl, libacehip), Cint,
                     (Ref{AceJlBasis}, Cint, Ptr{Float64}, Int64, Ref{Ptr{Cvoid}}),
                     jb, device, cv, length(cv), out))
    end
    ev = HIPEvaluator(out[], info.prop, info.nB, info.cat, device)
    finalizer(e -> (e.ctx != C_NULL && ccall((:ace_ctx_destroy, libacehip), Cvoid, (Ptr{Cvoid},), e.ctx);
                    e.ctx = C_NULL), ev)
    return ev
end

"""
    hip_model(basis, c = zeros(length(basis)); device = 0)

`LinearACEModel(basis, c, HIPEvaluator(basis, c))`.
"""
hip_model(basis::SymmetricBasis, c::AbstractVector = zeros(length(basis)); device::Integer = 0) =
    LinearACEModel(basis, Vector{Float64}(c), HIPEvaluator(basis, c; device = device))

# set_params!(m, c) -> set_params!(m.evaluator, m.basis, c) (src/linearmodel.jl:67-71)
function ACE.set_params!(ev::HIPEvaluator, basis::SymmetricBasis, c::AbstractVector)
    cv = Vector{Float64}(c)
    _check(ccall((:ace_set_params, libacehip), Cint, (Ptr{Cvoid}, Ptr{Float64}, Int64), ev.ctx, cv, length(cv)))
    return ev
end

# FIO: the evaluator's entry of a LinearACEModel file (src/linearmodel.jl:80-93)
ACE.write_dict(ev::HIPEvaluator) = Dict("__id__" => "ACEHIP_HIPEvaluator", "device" => ev.device)

ACE.read_dict(::Val{:ACEHIP_HIPEvaluator}, D::Dict, basis, c) =
    HIPEvaluator(basis, c; device = get(D, "device", 0))

# ---------------------------------------------------------------------------
# configurations -> ragged SoA batches

_states(cfg::ACE.ACEConfig) = cfg.Xs
_states(cfg::AbstractVector) = cfg

function _flatten(ev::HIPEvaluator, cfgs)
    nenv = length(cfgs)
    off = zeros(Int32, nenv + 1)
    for (i, cfg) in enumerate(cfgs)
        off[i+1] = off[i] + length(_states(cfg))
    end
    nnbr = Int(off[end])
    xyz = Matrix{Float64}(undef, nnbr, 3)          # column-major: the x, y, z blocks of xyz[3][nnbr]
    sp = ev.categorical === nothing ? Int8[] : Vector{Int8}(undef, nnbr)
    j = 0
    for cfg in cfgs, X in _states(cfg)
        j += 1
        xyz[j, 1], xyz[j, 2], xyz[j, 3] = X.rr[1], X.rr[2], X.rr[3]
        if ev.categorical !== nothing
            sp[j] = Int8(ACE.val2i(ev.categorical.categories, ACE._val(X, ev.categorical)) - 1)
        end
    end
    return off, xyz, sp
end

_sp(sp) = isempty(sp) ? C_NULL : pointer(sp)

function _energy_forces(ev::HIPEvaluator, cfgs; forces::Bool)
    off, xyz, sp = _flatten(ev, cfgs)
    nc = ev.prop == PROP_INVARIANT ? 1 : 3
    E = Matrix{Float64}(undef, nc, length(cfgs))
    g = Matrix{Float64}(undef, size(xyz, 1), 3 * nc)   # column q = p*3 + d of dE_dr[3 nc][nnbr]
    GC.@preserve off xyz sp E g begin
        if forces
            _check(ccall((:ace_energy_forces_host, libacehip), Cint,
                         (Ptr{Cvoid}, Int64, Ptr{Int32}, Ptr{Float64}, Ptr{Int8}, Ptr{Float64}, Ptr{Float64}),
                         ev.ctx, length(cfgs), off, xyz, _sp(sp), E, g))
        else
            _check(ccall((:ace_energy_host, libacehip), Cint,
                         (Ptr{Cvoid}, Int64, Ptr{Int32}, Ptr{Float64}, Ptr{Int8}, Ptr{Float64}),
                         ev.ctx, length(cfgs), off, xyz, _sp(sp), E))
        end
    end
    return E, g, off
end

_property(ev::HIPEvaluator, E, i) =
    ev.prop == PROP_INVARIANT ? ACE.Invariant(E[1, i]) : ACE.EuclideanVector(SVector{3}(E[1, i], E[2, i], E[3, i]))

"""
    evaluate_batch(m, ev::HIPEvaluator, cfgs)

Site energies (Invariant or EuclideanVector values) of every configuration
in `cfgs`, in one library call.
"""
function evaluate_batch(m::LinearACEModel, ev::HIPEvaluator, cfgs::AbstractVector)
    E, _, _ = _energy_forces(ev, cfgs; forces = false)
    return [_property(ev, E, i) for i = 1:length(cfgs)]
end

# the gradient entries of neighbour j: +dE/dr_j (Invariant), the Jacobian
# d phi_p / d r_j,d (EuclideanVector, as coco_o_daa builds it, properties.jl:44-55)
function _dstates(ev::HIPEvaluator, g, j0, j1)
    if ev.prop == PROP_INVARIANT
        return [DState(rr = SVector{3}(g[j, 1], g[j, 2], g[j, 3])) for j = j0:j1]
    end
    return [DState(rr = SMatrix{3,3}(ntuple(i -> g[j, 3 * ((i - 1) % 3) + (i - 1) ÷ 3 + 1], 9))) for j = j0:j1]
end

"""
    grad_config_batch(m, ev::HIPEvaluator, cfgs)

`grad_config` of every configuration in `cfgs`, in one library call.
"""
function grad_config_batch(m::LinearACEModel, ev::HIPEvaluator, cfgs::AbstractVector)
    _, g, off = _energy_forces(ev, cfgs; forces = true)
    return [_dstates(ev, g, off[i] + 1, off[i+1]) for i = 1:length(cfgs)]
end

# evaluate(m, m.evaluator, X) (src/linearmodel.jl:107-108; src/evaluator.jl:117-147)
ACE.evaluate(m::LinearACEModel, ev::HIPEvaluator, cfg::ACE.AbstractConfiguration) =
    evaluate_batch(m, ev, [cfg])[1]

# grad_config(m, m.evaluator, X) (src/linearmodel.jl:110-111; src/evaluator.jl:150-151)
ACE.grad_config(m::LinearACEModel, ev::HIPEvaluator, cfg::ACE.UConfig) =
    grad_config_batch(m, ev, [cfg])[1]

# _rrule_evaluate(dp, m, V, cfg) (src/evaluator.jl:154-200): the gradient of
# contract(dp, E) with respect to the neighbour positions
_dp_scalar(dp::Number) = dp
_dp_scalar(dp::ACE.Invariant) = dp.val
_dp_vector(dp::ACE.EuclideanVector) = dp.val
_dp_vector(dp::AbstractVector) = SVector{3}(dp[1], dp[2], dp[3])

function ACE._rrule_evaluate(dp, m::LinearACEModel, ev::HIPEvaluator, cfg::ACE.UConfig)
    _, g, _ = _energy_forces(ev, [cfg]; forces = true)
    n = size(g, 1)
    if ev.prop == PROP_INVARIANT
        s = _dp_scalar(dp)
        return [DState(rr = SVector{3}(s * g[j, 1], s * g[j, 2], s * g[j, 3])) for j = 1:n]
    end
    w = _dp_vector(dp)
    return [DState(rr = SVector{3}(ntuple(d -> w[1] * g[j, d] + w[2] * g[j, 3+d] + w[3] * g[j, 6+d], 3))) for j = 1:n]
end

# grad_params(m, cfg) = the basis values B (src/linearmodel.jl:114-123), on the GPU
function ACE.grad_params(m::LinearACEModel{<:Any,<:Any,HIPEvaluator}, cfg::ACE.AbstractConfiguration)
    ev = m.evaluator
    off, xyz, sp = _flatten(ev, [cfg])
    nc = ev.prop == PROP_INVARIANT ? 1 : 3
    B = Matrix{Float64}(undef, nc, ev.nB)
    GC.@preserve off xyz sp B begin
        _check(ccall((:ace_basis_B_host, libacehip), Cint,
                     (Ptr{Cvoid}, Int64, Ptr{Int32}, Ptr{Float64}, Ptr{Int8}, Ptr{Float64}),
                     ev.ctx, 1, off, xyz, _sp(sp), B))
    end
    if ev.prop == PROP_INVARIANT
        return [ACE.Invariant(B[1, k]) for k = 1:ev.nB]
    end
    # scalar coefficients: the basis values themselves (_gi(Bi, c::Number) = Bi)
    return [ACE.EuclideanVector(SVector{3}(B[1, k], B[2, k], B[3, k])) for k = 1:ev.nB]
end

# adjoint_EVAL_D(m, V, cfg, w) (src/evaluator.jl:204-244): dB_k = sum_j w_j . dB_k/dr_j
function ACE.adjoint_EVAL_D(m::LinearACEModel, ev::HIPEvaluator, cfg, w)
    ev.prop == PROP_INVARIANT || error("ACEHIP: adjoint_EVAL_D supports Invariant bases")
    off, xyz, sp = _flatten(ev, [cfg])
    n = size(xyz, 1)
    length(w) == n || error("ACEHIP: adjoint_EVAL_D needs one weight per neighbour")
    W = Matrix{Float64}(undef, n, 3)
    for j = 1:n
        W[j, 1], W[j, 2], W[j, 3] = w[j].rr[1], w[j].rr[2], w[j].rr[3]
    end
    dB = Vector{Float64}(undef, ev.nB)
    GC.@preserve off xyz sp W dB begin
        _check(ccall((:ace_adjoint_EVAL_D_host, libacehip), Cint,
                     (Ptr{Cvoid}, Int64, Ptr{Int32}, Ptr{Float64}, Ptr{Int8}, Ptr{Float64}, Ptr{Float64}),
                     ev.ctx, 1, off, xyz, _sp(sp), W, dB))
    end
    return [ACE.Invariant(x) for x in dB]
end

end # module
