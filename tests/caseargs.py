"""Decode a golden case's reference-driver options into oracle / library arguments."""
import numpy as np

from oracle import oracle as O

KIND = {"cg": O.CG, "bicgstab": O.BICGSTAB, "cheb": O.CHEB, "none": O.NONE}


def opt(args, key, n=1, default=None):
    if key not in args:
        return default
    i = args.index(key)
    vals = args[i + 1:i + 1 + n]
    return vals if n > 1 else vals[0]


def decode(case):
    a = case["args"]
    np_ = [int(v) for v in opt(a, "--np", 3)]
    ds = [float(v) for v in opt(a, "--ds", 3)]
    origin = [float(v) for v in opt(a, "--origin", 3, ["0", "0", "0"])]
    bc = [int(v) for v in opt(a, "--bc", 6, ["0"] * 6)]
    problem = int(opt(a, "--problem", 1, "0"))
    rhs = opt(a, "--rhs", 1, "function")
    seed = int(opt(a, "--seed", 1, "1234"))
    speck = int(opt(a, "--spec-k", 1, "16"))
    solver = opt(a, "--solver", 1, "cg")
    precond = opt(a, "--precond", 1, "none")
    m = int(opt(a, "--cheb-m", 1, "11"))
    eps = float(opt(a, "--cheb-eps", 1, "1e-4"))
    resc = [float(v) for v in opt(a, "--cheb-rescale", 2, ["500", "0.9999"])]
    tol = case.get("tol", 1e-10)
    chain = [dict(kind=KIND[solver], tol=tol, maxit=m if solver == "cheb" else 100000,
                  cheb_eps=eps, cheb_rmin=resc[0], cheb_rmax=resc[1])]
    if solver != "cheb":
        if precond == "cheb":
            chain.append(dict(kind=O.CHEB, maxit=m, cheb_eps=eps, cheb_rmin=resc[0], cheb_rmax=resc[1]))
        elif precond in ("bicgstab", "cg"):
            chain += [dict(kind=KIND[precond], tol=10000 * tol, maxit=150), dict(kind=O.NONE)]
        else:
            chain.append(dict(kind=O.NONE))
    return dict(np=np_, ds=ds, origin=origin, bc=bc, problem=problem, rhs=rhs, seed=seed, speck=speck, chain=chain)


def oracle_problem(d):
    """Oracle grid + x, b, flux for a decoded case (b recipe as the reference driver)."""
    G = O.OracleGrid(d["np"], d["ds"], d["origin"], d["bc"])
    x, b, flux = G.set_problem(d["problem"])
    sl = G.solver_slice()
    if d["rhs"] in ("normal", "gauss", "spectrum", "multiscale"):
        if d["rhs"] == "normal":
            b[sl] = G.rhs("normal", d["seed"])[sl]
        elif d["rhs"] == "gauss":
            b[sl] = G.rhs("gauss", d["seed"])[sl]
        elif d["rhs"] == "spectrum":
            b[sl] = G.rhs("spectrum", d["seed"], K=d["speck"])[sl]
        else:
            b[sl] = G.rhs("gauss", d["seed"])[sl] + G.rhs("spectrum", d["seed"], K=d["speck"])[sl]
    return G, x, b, flux


def data_region(G, x):
    L = G.lim_data
    return x[L[4]:L[5], L[2]:L[3], L[0]:L[1]]


def relmax(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
