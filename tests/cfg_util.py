"""Test helper: the oracle problem matching a workload of solidfem_amd.configs."""
from oracle import oracle as O
from solidfem_amd.mesh import lame


def oracle_problem(d, nthreads=0, modulus_scale=1.0):
    """modulus_scale multiplies the Lame parameters (a relative perturbation of
    the operator, for sensitivity checks)."""
    m = d["mesh"]
    if d["young_e"] is not None:
        mu, lam = lame(d["young_e"], d["poisson_e"])
        kw = dict(mu=mu * modulus_scale, lam=lam * modulus_scale)
    elif modulus_scale != 1.0:
        mu, lam = lame(d["young"], d["poisson"])
        kw = dict(mu=float(mu) * modulus_scale, lam=float(lam) * modulus_scale)
    else:
        kw = dict(young=d["young"], poisson=d["poisson"])
    return O.OracleProblem(m.X, m.tets, d["fixed"], d["material"], density=1070.0, gravity=d["gravity"],
                           tris=d["tris"], pressure=d["pressure"], fext=d["fdead"], nthreads=nthreads,
                           incidence=True, **kw)
