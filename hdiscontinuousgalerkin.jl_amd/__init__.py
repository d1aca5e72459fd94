"""MI355X HDG Poisson pipeline: the hot path of HDiscontinuousGalerkin.jl's
HDG driver (examples/poisson2D_HDG.jl) as libhdg (HIP, gfx950) plus a host
mirror of the reference's API.

    import hdg_amd as H           # (this directory, loaded under any name)
    mesh = H.rectangle_mesh((32, 32))
    fe = H.GenericFiniteElement(H.Dubiner(1))
    Wh = H.ScalarFunctionSpace(mesh, fe); Vh = H.VectorFunctionSpace(mesh, fe)
    Mh = H.ScalarTraceFunctionSpace(Wh, H.GenericFiniteElement(H.Legendre(1)))
    uh_hat, sigma_h, u_h = H.TrialFunction(Mh), H.TrialFunction(Vh), H.TrialFunction(Wh)
    dbc = H.Dirichlet(uh_hat, mesh, "boundary", lambda x: 0.0)
    H.hdg_solve(sigma_h, u_h, uh_hat, mesh, dbc)
    H.errornorm(u_h, "sinsin")
"""
from ._lib import (ArgumentError, HDGError, SingularException, HDG_FP64, HDG_MIXED, lib,  # noqa: F401
                   LIB_PATH, EXPORTS, PC_AUTO, PC_BLOCK_JACOBI, PC_MULTIGRID, SCHEDULE_EAGER,
                   SCHEDULE_AUTOTUNE, SCHEDULE_LU)
from .context import Context  # noqa: F401
from .server import Batch, Server, footprint, make_request  # noqa: F401
from .hdg import (apply_bc, default_context, doassemble, errornorm, get_usigma, hdg_solve,  # noqa: F401
                  meandiag)
from .mesh import (PolygonalMesh, cell_volume, face_orientation, get_vertices_matrix,  # noqa: F401
                   getcells_matrix, getfaceset, getncells, getnfaces, getnnodes, getnodeset,
                   n_faces_per_cell, parse_mesh_triangle, rectangle_mesh)
from .spaces import (Dirichlet, Dubiner, GenericFiniteElement, Legendre, ScalarFunctionSpace,  # noqa: F401
                     ScalarTraceFunctionSpace, TrialFunction, VectorFunctionSpace, getnbasefunctions,
                     getnlocaldofs, quadrature)
