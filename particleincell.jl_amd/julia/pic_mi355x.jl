# pic_mi355x.jl -- the Julia side of libpicmi355x, for ParticleInCell.jl.
#
# Two submodules, included from ParticleInCell.jl's module next to
# src/pic.jl (`include("pic_mi355x.jl")`):
#
#   F90     the drop-in of src/pic.jl:2-116: the same methods, ccall'ing the
#           library's pic_f90_* symbols, which take src/pic.f90's arguments
#           (scalars by reference, column-major host arrays) and return a
#           status code instead of nothing.
#   MI355X  device-resident methods of the reference's generic functions on a
#           DeviceParticles handle (a pic_ctx): compute_rho, push_v!, push_x!,
#           solve!, the Vlasov-Maxwell calls of src/fdtd.jl, the UA step of
#           src/ua_steps.jl, the device quiet start and read_particles, and
#           the fused time loops.
#
# The library is found through ENV["PICMI355X_LIB"] (a path) or the loader's
# search path.  Every call checks the returned status and raises the
# library's message (pic_last_error).  C signatures: include/pic_mi355x.h.

module F90

using Libdl
import ..ParticleInCell: TwoDGrid, CloudInCell

const piclib = get(ENV, "PICMI355X_LIB", "libpicmi355x")  # src/pic.jl:6 builds libpic

check(rc) = rc == 0 || error(unsafe_string(ccall((:pic_last_error, piclib), Cstring, ())))

# src/pic.jl:18-57: the CIC current of the Fortran `deposition`, ghosts folded
function compute_current!(jx, jy, m::TwoDGrid, kernel::CloudInCell, p)
    check(ccall((:pic_f90_deposition, piclib), Cint,
                (Ref{Int32}, Ref{Int32}, Ref{Int32}, Ref{Float64}, Ref{Float64},
                 Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                Int32(size(p.array, 2)), Int32(m.nx), Int32(m.ny), m.dx, m.dy,
                p.array, jx, jy))
    nx, ny = m.nx, m.ny
    for i = 1:nx+1
        jx[i, 1] += jx[i, ny+1]
        jx[i, ny+1] = jx[i, 1]
    end
    for j = 1:ny+1
        jx[1, j] += jx[nx+1, j]
        jx[nx+1, j] = jx[1, j]
    end
end

# src/pic.jl:60-75
function push_x!(p, mesh::TwoDGrid, dt::Float64)
    check(ccall((:pic_f90_push_x, piclib), Cint,
                (Ref{Int32}, Ref{Float64}, Ref{Float64}, Ptr{Float64}, Ref{Float64}),
                Int32(size(p.array, 2)), mesh.dimx, mesh.dimy, p.array, dt))
end

# src/pic.jl:78-114 (fields (nx+1) x (ny+1), as the Fortran twin reads them)
function push_v!(p, kernel::CloudInCell, m::TwoDGrid, ex, ey, bz, dt::Float64)
    check(ccall((:pic_f90_push_v, piclib), Cint,
                (Ref{Int32}, Ptr{Float64}, Ref{Float64}, Ref{Int32}, Ref{Int32},
                 Ref{Float64}, Ref{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                Int32(size(p.array, 2)), p.array, dt, Int32(m.nx), Int32(m.ny), m.dx, m.dy,
                ex, ey, bz))
end

end  # module F90

module MI355X

using Libdl
import ..ParticleInCell: TwoDGrid, CloudInCell, M6, NearestGridPoint, TriangularShapeCloud, M4,
                        compute_rho, push_v!, push_x!, solve!
import ..ParticleInCell: FDTD, faraday!, ampere_maxwell!, update_fields!, compute_current!,
                        compute_energy
import ..ParticleInCell: UA, Particles, preparation!, compute_f!, ua_step!

const lib = get(ENV, "PICMI355X_LIB", "libpicmi355x")

const PIC_CIC, PIC_M6, PIC_NGP, PIC_TSC, PIC_M4 = Cint(0), Cint(1), Cint(2), Cint(3), Cint(4)
const PIC_GRID_RHO, PIC_GRID_EX, PIC_GRID_EY, PIC_GRID_BZ = Cint(0), Cint(1), Cint(2), Cint(3)
const PIC_BZ_NONE, PIC_BZ_UNIFORM, PIC_BZ_GRID = Cint(0), Cint(1), Cint(2)

check(rc) = rc == 0 || error(unsafe_string(ccall((:pic_last_error, lib), Cstring, ())))

code(::CloudInCell) = PIC_CIC
code(::M6) = PIC_M6
code(::NearestGridPoint) = PIC_NGP
code(::TriangularShapeCloud) = PIC_TSC
code(::M4) = PIC_M4

"Particles and grids of one mesh resident on the GPU (a pic_ctx)."
mutable struct DeviceParticles
    h::Ptr{Cvoid}
    mesh::TwoDGrid
    n::Int
end

function DeviceParticles(mesh::TwoDGrid, capacity::Integer; device = 0)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:pic_ctx_create, lib), Cint,
                (Ref{Ptr{Cvoid}}, Cint, Cint, Cint, Float64, Float64, Float64, Float64, Int64),
                h, device, mesh.nx, mesh.ny, mesh.xmin, mesh.xmax, mesh.ymin, mesh.ymax,
                capacity))
    d = DeviceParticles(h[], mesh, 0)
    finalizer(x -> ccall((:pic_ctx_destroy, lib), Cint, (Ptr{Cvoid},), x.h), d)
    d
end

"A ParticleGroup's p.array (5 x n: x, y, vx, vy, w) copied to the GPU."
function DeviceParticles(array::Matrix{Float64}, mesh::TwoDGrid; device = 0)
    d = DeviceParticles(mesh, size(array, 2); device = device)
    check(ccall((:pic_set_particles, lib), Cint, (Ptr{Cvoid}, Ptr{Float64}, Int64),
                d.h, array, size(array, 2)))
    d.n = size(array, 2)
    d
end

function Base.copyto!(array::Matrix{Float64}, d::DeviceParticles)   # p.array <- device
    check(ccall((:pic_get_particles, lib), Cint, (Ptr{Cvoid}, Ptr{Float64}), d.h, array))
    array
end

"FP64 (false) or FP32 cell-relative (true) particle state on the device."
set_fp32!(d::DeviceParticles, on::Bool) =
    check(ccall((:pic_set_state, lib), Cint, (Ptr{Cvoid}, Cint), d.h, on ? 1 : 0))

function set_grid!(d::DeviceParticles, which, a)
    check(ccall((:pic_set_grid, lib), Cint, (Ptr{Cvoid}, Cint, Ptr{Float64}, Cint, Cint),
                d.h, which, a, size(a, 1), 1))
end

"Grid `which` as nx x ny, or (nx+1) x (ny+1) with the periodic ghosts."
function get_grid(d::DeviceParticles, which; ghost = false)
    nx, ny = d.mesh.nx, d.mesh.ny
    a = ghost ? zeros(nx + 1, ny + 1) : zeros(nx, ny)
    check(ccall((:pic_get_grid, lib), Cint, (Ptr{Cvoid}, Cint, Ptr{Float64}, Cint, Cint, Cint),
                d.h, which, a, size(a, 1), 1, ghost ? 1 : 0))
    a
end

function set_bz!(d::DeviceParticles, bz)
    if bz isa Number
        check(ccall((:pic_set_bz, lib), Cint, (Ptr{Cvoid}, Cint, Float64), d.h,
                    bz == 0 ? PIC_BZ_NONE : PIC_BZ_UNIFORM, bz))
    else
        set_grid!(d, PIC_GRID_BZ, bz)
        check(ccall((:pic_set_bz, lib), Cint, (Ptr{Cvoid}, Cint, Float64), d.h, PIC_BZ_GRID, 0.0))
    end
end

# compute_rho(p, kernel, mesh): pushers_cic.jl:75-114, pushers_m6.jl:226-348
function compute_rho(d::DeviceParticles, kernel, mesh::TwoDGrid)
    check(ccall((:pic_compute_rho, lib), Cint, (Ptr{Cvoid}, Cint), d.h, code(kernel)))
    get_grid(d, PIC_GRID_RHO)
end

# push_v!(p, kernel, m, ex, ey, bz, dt): pushers_cic.jl:3-53, pushers_m6.jl:5-224
function push_v!(d::DeviceParticles, kernel, m::TwoDGrid, ex, ey, bz, dt)
    set_grid!(d, PIC_GRID_EX, ex)
    set_grid!(d, PIC_GRID_EY, ey)
    set_bz!(d, bz)
    check(ccall((:pic_push_v, lib), Cint, (Ptr{Cvoid}, Cint, Float64), d.h, code(kernel), dt))
end

# push_x!(p, mesh, dt): pushers_cic.jl:55-72
push_x!(d::DeviceParticles, mesh::TwoDGrid, dt::Float64) =
    check(ccall((:pic_push_x, lib), Cint, (Ptr{Cvoid}, Float64), d.h, dt))

# solve!(ex, ey, poisson, rho): poisson_2d.jl:97-113, on the device's rho
function solve!(ex, ey, d::DeviceParticles)
    check(ccall((:pic_solve_e, lib), Cint, (Ptr{Cvoid},), d.h))
    ex .= get_grid(d, PIC_GRID_EX)
    ey .= get_grid(d, PIC_GRID_EY)
end

"""
The hot-path driver's loop (examples/landau_damping_poisson_cic_m6.jl:70-83)
fused on the device: nsteps iterations, the particles lane-sorted every
sort_every steps; returns sqrt(sum(ex^2) dx dy) per step.
"""
function run!(d::DeviceParticles, kernel, dt, nsteps; sort_every = 8)
    e = zeros(2, nsteps)
    check(ccall((:pic_run, lib), Cint, (Ptr{Cvoid}, Cint, Float64, Cint, Cint, Ptr{Float64}),
                d.h, code(kernel), dt, nsteps, sort_every, e))
    sqrt.(e[1, :])
end

# ---------------------------------------------------- Vlasov-Maxwell (fdtd.jl)

fdtd_in!(d::DeviceParticles, f::FDTD) = check(ccall((:pic_fdtd_set, lib), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}), d.h, f.ex, f.ey, f.bz))
fdtd_out!(f::FDTD, d::DeviceParticles) = check(ccall((:pic_fdtd_get, lib), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}), d.h, f.ex, f.ey, f.bz))

function faraday!(f::FDTD, d::DeviceParticles, dt)                   # fdtd.jl:24
    fdtd_in!(d, f)
    check(ccall((:pic_faraday, lib), Cint, (Ptr{Cvoid}, Float64), d.h, dt))
    fdtd_out!(f, d)
end

function ampere_maxwell!(f::FDTD, d::DeviceParticles, jx, jy, dt)    # fdtd.jl:39
    fdtd_in!(d, f)
    check(ccall((:pic_set_current, lib), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}),
                d.h, jx, jy))
    check(ccall((:pic_ampere_maxwell, lib), Cint, (Ptr{Cvoid}, Float64), d.h, dt))
    fdtd_out!(f, d)
end

function update_fields!(ex, ey, bz, d::DeviceParticles, f::FDTD)     # fdtd.jl:58
    fdtd_in!(d, f)
    check(ccall((:pic_update_fields, lib), Cint, (Ptr{Cvoid},), d.h))
    for (a, w) in ((ex, PIC_GRID_EX), (ey, PIC_GRID_EY), (bz, PIC_GRID_BZ))
        check(ccall((:pic_get_grid, lib), Cint, (Ptr{Cvoid}, Cint, Ptr{Float64}, Cint, Cint, Cint),
                    d.h, w, a, size(a, 1), 1, 1))
    end
end

compute_current!(jx, jy, d::DeviceParticles, ::CloudInCell) =        # cloud_in_cell.jl:10
    check(ccall((:pic_compute_current, lib), Cint, (Ptr{Cvoid}, Cint, Ptr{Float64}, Ptr{Float64}),
                d.h, PIC_CIC, jx, jy))

function compute_energy(f::FDTD, d::DeviceParticles)                 # fdtd.jl:79
    fdtd_in!(d, f)
    e = Ref(0.0)
    check(ccall((:pic_fdtd_energy, lib), Cint, (Ptr{Cvoid}, Ref{Float64}), d.h, e))
    e[]
end

"examples/vm2d.jl:40-56 on the device; f holds the fields after the last step."
function run_maxwell!(d::DeviceParticles, f::FDTD, dt, nsteps)
    fdtd_in!(d, f)
    e = zeros(nsteps)
    check(ccall((:pic_run_maxwell, lib), Cint, (Ptr{Cvoid}, Cint, Float64, Cint, Ptr{Float64}),
                d.h, PIC_CIC, dt, nsteps, e))
    fdtd_out!(f, d)
    e
end

# ------------------------------------------------------------ UA (ua_steps.jl)

const CPtr = Ptr{ComplexF64}

function bind!(d::DeviceParticles, ua::UA, p::Particles)
    check(ccall((:pic_set_particles_xv, lib), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64},
                Float64, Int64), d.h, p.x, p.v, p.w, p.nbpart))
    check(ccall((:pic_set_particle_fields, lib), Cint, (Ptr{Cvoid}, Ptr{Float64}), d.h, p.e))
    check(ccall((:pic_ua_init, lib), Cint, (Ptr{Cvoid}, Cint, Float64), d.h, ua.ntau, ua.ε))
    d.n = p.nbpart
end

function preparation!(ua::UA, dt, p::Particles, xt, yt, d::DeviceParticles)  # ua_steps.jl:3
    bind!(d, ua, p)
    check(ccall((:pic_ua_preparation, lib), Cint,
                (Ptr{Cvoid}, Float64, Ptr{Float64}, Ptr{Float64}, CPtr, CPtr, CPtr, CPtr),
                d.h, dt, p.b, p.t, ua.pl, ua.ql, xt, yt))
end

compute_f!(fx, fy, ua::UA, p::Particles, xt, yt, et, d::DeviceParticles) =   # ua_steps.jl:107
    check(ccall((:pic_ua_compute_f, lib), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, CPtr, CPtr, Ptr{Float64}, CPtr, CPtr),
                d.h, p.b, xt, yt, et, fx, fy))

ua_step!(xt, x̃t, ua::UA, p::Particles, fx, d::DeviceParticles) =            # ua_steps.jl:152
    check(ccall((:pic_ua_step, lib), Cint, (Ptr{Cvoid}, Ptr{Float64}, CPtr, CPtr, CPtr, CPtr),
                d.h, p.t, x̃t, fx, C_NULL, xt))

ua_step!(xt, x̃t, ua::UA, p::Particles, fx, gx, d::DeviceParticles) =        # ua_steps.jl:174
    check(ccall((:pic_ua_step, lib), Cint, (Ptr{Cvoid}, Ptr{Float64}, CPtr, CPtr, CPtr, CPtr),
                d.h, p.t, x̃t, fx, gx, xt))

"test/bupdate.jl:62-116 on the device after its prologue; the two solve! energies per step."
function ua_run!(d::DeviceParticles, ua::UA, dt, nsteps)
    check(ccall((:pic_ua_init, lib), Cint, (Ptr{Cvoid}, Cint, Float64), d.h, ua.ntau, ua.ε))
    e = zeros(2, nsteps)
    check(ccall((:pic_ua_run, lib), Cint, (Ptr{Cvoid}, Float64, Cint, Ptr{Float64}),
                d.h, dt, nsteps, e))
    e
end

# ----------------------------------------------- initial conditions and files

"sample!(pg, mesh, LandauDamping(alpha, kx)) (landau_damping.jl:77-97) on the device."
function sample_landau!(d::DeviceParticles, n, alpha, kx; fp32 = false)
    check(ccall((:pic_sample_landau, lib), Cint, (Ptr{Cvoid}, Int64, Float64, Float64, Cint),
                d.h, n, alpha, kx, fp32 ? 1 : 0))
    d.n = n
end

"Density 1 + alpha cos(kx x) cos(ky y), Maxwellian velocities of width vth, seeded."
function sample_maxwellian!(d::DeviceParticles, n, alpha, kx, ky; vth = 1.0, seed = 0,
                            fp32 = false)
    check(ccall((:pic_sample_maxwellian, lib), Cint,
                (Ptr{Cvoid}, Int64, Float64, Float64, Float64, Float64, UInt64, Cint),
                d.h, n, alpha, kx, ky, vth, seed, fp32 ? 1 : 0))
    d.n = n
end

"read_particles(filename, mesh) (read_particles.jl:3-35), parsed by the library."
function read_particles!(d::DeviceParticles, filename; fp32 = false)
    n = Ref{Int64}(0)
    check(ccall((:pic_read_particles, lib), Cint, (Ptr{Cvoid}, Cstring, Cint, Ref{Int64}),
                d.h, filename, fp32 ? 1 : 0, n))
    d.n = n[]
end

"""
compute_rho(p, kernel, mesh) of a p.array already in device memory into a
device rho, allocating nothing (for callers under HIP graph capture or that
forbid hipMalloc); `stream` is a hipStream_t (C_NULL: the null stream).
"""
compute_rho_noalloc!(d_rho::Ptr{Float64}, d_p::Ptr{Float64}, n, kernel, mesh::TwoDGrid;
                     stream = C_NULL) =
    check(ccall((:pic_compute_rho_noalloc, lib), Cint,
                (Cint, Ptr{Float64}, Int64, Cint, Cint, Float64, Float64, Float64, Float64,
                 Ptr{Float64}, Ptr{Cvoid}),
                code(kernel), d_p, n, mesh.nx, mesh.ny, mesh.xmin, mesh.xmax, mesh.ymin,
                mesh.ymax, d_rho, stream))

end  # module MI355X
