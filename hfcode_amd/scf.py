"""Direct RHF SCF driver around a Fock builder.

A restatement of the reference driver ``direct_hartreefock``
(direct_hartreefock.f90:1-234) for jobs that run without the Fortran front
end; the Fock builds are the drop-in (:class:`hfcode_amd.fock.FockBuilder`)
or anything with the same four methods.  Steps, as the reference takes them:

* S^-1/2 by symmetric orthogonalisation (:24-29), core-Hamiltonian guess (:66-72);
* iteration i (:98-223): full build (initial_diag + non_diag at i = 1,
  diag + non_diag when a full rebuild is due) or the incremental build on
  dP = P - P_old with the non-diagonal threshold x0.1, full rebuild forced
  every 15 incremental steps (:100-121); with a mixed-precision builder the
  incremental builds still run in FP64;
* E_elec = 1/2 sum P (H + F) (:123-131); DIIS error S^-1/2 (FPS - SPF) S^-1/2
  (:133-144); DIIS over up to 8 Fock matrices with 1.05 damping of all but
  the lowest-energy entry and a 1.5 coefficient cap (:161-199);
* threshold switch 5e-7 -> 1e-10 with a full rebuild and DIIS reset when
  i >= 7 and max|resid| < 0.005 (:152-160);
* optional level shift of the virtual orbitals by `levelshift` hartree in
  the basis of the previous iteration's orbitals while no DIIS runs (:39-46,
  :200-203);
* convergence: |dE| < 5e-8 and max|dP| < 1e-5 (:207-216), at most 25
  iterations; non-convergence is reported (the reference is silent).

Linear algebra is FP64 numpy (eigh for dsyev, solve for dsysv).
"""
import time

import numpy as np

MAXITER = 25
NDIIS = 8
CONV_P = 1e-5
CONV_E = 5e-8
LOWTHRESH = 5e-7
INTTHRESH = 1e-10
FDIFFFAC = 0.1


def core_guess(S, H, nocc):
    """Core-Hamiltonian guess density, as direct_hartreefock.f90:24-29,66-72 forms it."""
    w, U = np.linalg.eigh(S)
    X = (U * (1.0 / np.sqrt(w))) @ U.T
    _, C = np.linalg.eigh(X @ H @ X.T)
    C = X @ C
    return 2.0 * C[:, :nocc] @ C[:, :nocc].T


class SCFResult:
    def __init__(self, **kw):
        self.__dict__.update(kw)


def rhf(builder, S, H, nocc, unuc, maxiter=MAXITER, ndiis=NDIIS, conv_p=CONV_P, conv_e=CONV_E,
        lowthresh=LOWTHRESH, intthresh=INTTHRESH, P0=None, levelshift=None, log=None, diis_start=0.3):
    """Run the direct SCF; returns an SCFResult (energy, P, F, iterations, converged).

    `diis_start`: max|DIIS error| below which DIIS starts and the level shift
    stops (0.3 in the reference, direct_hartreefock.f90:146-150)."""
    nb = S.shape[0]
    w, U = np.linalg.eigh(S)
    X = (U * (1.0 / np.sqrt(w))) @ U.T                     # S^-1/2
    resido = np.full((ndiis, nb, nb), 1000.0)
    fold = np.zeros((ndiis, nb, nb))
    eo = np.zeros(ndiis)
    diis = False
    fullfock = True
    thresh = lowthresh
    cmatob = None
    if P0 is None:
        fob = X @ H @ X.T
        _, cmatob = np.linalg.eigh(fob)
        cmat = X @ cmatob
        p = 2.0 * cmat[:, :nocc] @ cmat[:, :nocc].T
        uselevshft = True
    else:
        p = np.array(P0, dtype=np.float64)
        uselevshft = False
    shft = np.zeros(nb)
    if levelshift:
        shft[nocc:] = levelshift
    pold = p.copy()
    # E of the guess: sum P H for the core guess (direct_hartreefock.f90:83-88), 0 for a given density
    eelec = float(np.sum(p * H)) if P0 is None else 0.0
    nusediis = 0
    nfull = 0
    faobold = None
    converged = False
    history = []
    etot = None
    t_fock = 0.0
    for it in range(1, maxiter + 1):
        eeold = eelec
        tf0 = time.perf_counter()
        if fullfock and it == 1:
            f = builder.initial_diag(H, p, thresh)
            f = builder.non_diag(f, p, thresh)
            nfull = 0
        elif fullfock:
            f = builder.diag(H, p, thresh)
            f = builder.non_diag(f, p, thresh)
            fullfock = False
            nfull = 0
        else:
            delp = p - pold
            z = np.zeros((nb, nb))
            delf = builder.diag(z, delp, thresh)
            # FP64 whatever the builder's mixed split (an absolute bound: near
            # convergence almost every quartet of dP would fall below it)
            if getattr(builder, "split", 0.0):
                delf = builder.non_diag(delf, delp, thresh * FDIFFFAC, fp64=True)
            else:
                delf = builder.non_diag(delf, delp, thresh * FDIFFFAC)
            f = faobold + delf
            nfull += 1
            if nfull == 15:
                fullfock = True
        t_fock += time.perf_counter() - tf0
        eelec = 0.5 * float(np.sum(p * (H + f)))
        deltae = eelec - eeold
        etot = eelec + unuc
        faobold = f
        fob = X @ f @ X.T
        comm = f @ p @ S - S @ p @ f
        resid = X.T @ comm @ X
        resido[1:] = resido[:-1].copy()
        fold[1:] = fold[:-1].copy()
        eo[1:] = eo[:-1].copy()
        resido[0] = resid
        resmax = float(np.abs(resid).max())
        fold[0] = fob
        eo[0] = eelec
        if it >= 2 and resmax < diis_start and not diis:
            diis = True
            uselevshft = False
        if it >= 7 and thresh > intthresh and resmax < 0.005:
            eo[:] = 2000.0
            nusediis = 0
            fullfock = True
            uselevshft = False
            thresh = intthresh
        if diis:
            nusediis = min(nusediis + 1, ndiis)
            if nusediis > 1:
                n = nusediis
                B = np.zeros((n + 1, n + 1))
                B[n, :] = -1.0
                B[:, n] = -1.0
                B[n, n] = 0.0
                rhs = np.zeros(n + 1)
                rhs[n] = -1.0
                for j in range(n):
                    for k in range(j + 1):
                        B[j, k] = B[k, j] = float(np.sum(resido[j] * resido[k]))
                ii = int(np.argmin(eo))
                for j in range(n):
                    if j != ii:
                        B[j, j] *= 1.05
                sol = np.linalg.solve(B, rhs)
                if np.abs(sol[:n]).max() < 1.5:
                    fob = np.tensordot(sol[:n], fold[:n], axes=1)
                else:
                    nusediis = 0
        if uselevshft and levelshift and cmatob is not None:
            fob = fob + (cmatob * shft) @ cmatob.T
        eigs, cmatob = np.linalg.eigh(fob)
        cmat = X @ cmatob
        deltap = float(np.abs(p - pold).max())
        history.append((it, etot, deltae, deltap))
        if log:
            log("%3d %20.10f %20.10f %20.10f" % (it, etot, deltae, deltap))
        if abs(deltae) < conv_e and deltap < conv_p and it > 1:
            converged = True
            break
        pold = p
        p = 2.0 * cmat[:, :nocc] @ cmat[:, :nocc].T
    return SCFResult(energy=etot, P=p, F=faobold, eigs=eigs, iterations=it, converged=converged,
                     history=history, t_fock=t_fock)
