"""Host-side mirror of the reference plugin entry (ugacc.cc:54-117) over the
gfx950 library.

read_options() returns the reference's option names and defaults
(ugacc.cc:58-68).  ugacc(ref, options) performs the same checks with the
same messages (ugacc.cc:95-99), builds the CC wave function and runs
compute_energy (ugacc.cc:109-117), returning the reference wave function as
the reference does for DERTYPE=NONE.  CCWfn.compute_energy mirrors
ccwfn.cc:188-258: iteration table, summary, (T) for WFN=CCSD_T (and the
(T) gradient terms with DERTYPE=FIRST, :252-255) and "Computation has not
converged." on failure.  DERTYPE=FIRST then runs HBAR, Lambda and the
densities and prints the density energies (ugacc.cc:118-141), setting
ref.current_energy.  DERTYPE=SECOND adds the linear response of
ugacc.cc:143-241: the dipole ("Mu") and magnetic-dipole ("L") perturbed wave
functions per Cartesian component at MY_OMEGA (right-hand; with MYHAND=LEFT
also left-hand; with MYHAND=RIGHT and MY_OMEGA != 0 also Mu at -MY_OMEGA),
the dipole polarizabilities <<Mu_p;Mu_q>> and the length-gauge optical
rotation tensor <<Mu_p;L_q>>, printed and stored as ref.polarizabilities and
ref.rotations (label -> value, the reference's labels).

``ref`` is any object with the attributes of the front end's
chem.scf.Reference: no, nv, fock (active Fock), blocks (six Dirac blocks),
escf, and optionally enuc, efzc, reference ("RHF"), socc (open-shell count);
DERTYPE=SECOND also needs perturbation(op) -> three nact x nact MO matrices
(Perturbation::prop_p, perturbation.cc).
"""
import sys

from . import ugx


class PsiException(RuntimeError):
    """The reference's PSIEXCEPTION."""


def read_options(name="UGACC"):
    """Option names and defaults of read_options (ugacc.cc:58-68)."""
    return {
        "PRINT": 1, "REFERENCE": "RHF", "WFN": "CCSD", "DERTYPE": "NONE", "MYHAND": "RIGHT",
        "MAXITER": 100, "DIIS": True, "R_CONVERGENCE": 1e-7, "MY_OMEGA": 0.0, "OOC": False,
        "PROPERTY": "POLARIZABILITY",
        # library forms (no counterpart in the reference)
        "PRECISION": "FP64", "POLISH_THRESHOLD": 1e-5, "ENERGY_FORM": "REFERENCE",
    }


class CCWfn:
    """ccwfn.h's CCWfn over a device context (ccwfn.cc:12-120 + :188-258)."""

    def __init__(self, ref, options, device=0, out=sys.stdout):
        self.ref = ref
        self.options = options
        self.out = out
        self.wfn = options["WFN"]
        if self.wfn not in ("CCSD", "CCSD_T"):
            raise PsiException("Unknown WFN: %s" % self.wfn)
        self.ctx = ugx.Context(device)
        self.ctx.load_hamiltonian(ref.no, ref.nv, ref.fock, ref.blocks)
        self.result = None

    def _opts(self):
        o = ugx.default_options()
        opt = self.options
        o.wfn_ccsd_t = 1 if self.wfn == "CCSD_T" else 0
        o.maxiter = int(opt["MAXITER"])
        o.diis = 1 if opt["DIIS"] else 0
        o.r_convergence = float(opt["R_CONVERGENCE"])
        o.ooc = 1 if opt["OOC"] else 0
        o.precision = ugx.PREC_NAMES[str(opt.get("PRECISION", "FP64")).lower()]
        o.polish_threshold = float(opt.get("POLISH_THRESHOLD", 1e-5))
        o.energy_form = ugx.ENERGY_FULL if str(opt.get("ENERGY_FORM", "REFERENCE")).upper() == "FULL" \
            else ugx.ENERGY_REFERENCE
        o.print = 1 if int(opt["PRINT"]) > 0 else 0
        return o

    def compute_energy(self):
        r = self.ctx.ccsd(self._opts())
        self.result = r
        if not r["converged"]:
            raise PsiException("Computation has not converged.")
        eref = self.ref.escf
        w = self.out.write
        w("\n\tReference Energy           = %20.14f\n" % eref)
        w("\tMP2 Correlation Energy     = %20.14f\n" % r["e_mp2"])
        w("\tMP2 Total Energy           = %20.14f\n" % (r["e_mp2"] + eref))
        w("\tCCSD Correlation Energy    = %20.14f\n" % r["e_ccsd"])
        w("\tCCSD Total Energy          = %20.14f\n" % (r["e_ccsd"] + eref))
        etotal = r["e_ccsd"]
        if self.wfn == "CCSD_T":
            w("\t(T) Correction             = %20.14f%s\n" % (r["e_t"], " (TJL)" if self.options["OOC"] else ""))
            w("\tCCSD(T) Correlation Energy = %20.14f\n" % (r["e_ccsd"] + r["e_t"]))
            w("\tCCSD(T) Total Energy       = %20.14f\n" % (eref + r["e_ccsd"] + r["e_t"]))
            etotal += r["e_t"]
        if self.wfn == "CCSD_T" and self.options.get("DERTYPE", "NONE") == "FIRST":
            # tgrad, or tgrad_ooc with OOC=true (ccwfn.cc:252-255)
            self.ctx.tgrad(which=(), ooc=bool(self.options["OOC"]))
        return etotal

    def close(self):
        self.ctx.close()


CART = ("X", "Y", "Z")


def _std_to_string(x):
    """std::to_string of a double (the "%f" format), used in the reference's
    perturbation keys."""
    return "%f" % x


def _solve_pert(cc, g, hand, opts):
    """CCPert::solve (ccpert.cc:177-228): the banner, the converged line, and
    "Computation has not converged." when the residual stays above
    R_CONVERGENCE."""
    w = cc.out.write
    w("\n\tCoupled-Cluster %s Perturbed Wfn Iteration:\n" % ("RH" if hand == 0 else "LH"))
    w("\t-------------------------------------\n")
    w("\t Iter   Pseudoresponse      RMS   \n")
    w("\t-------------------------------------\n")
    r = g.solve(hand, opts)
    w("\t  %3d  %20.15f  %5.3e\n" % (r["niter"], r["pseudoresponse"], r["rms"]))
    if not r["converged"]:
        raise PsiException("Computation has not converged.")
    return r


def _linear_response(cc, ref, options):
    """ugacc.cc:143-241: perturbed wave functions, polarizabilities and the
    length-gauge optical rotation tensor."""
    if not hasattr(ref, "perturbation"):
        raise PsiException("DERTYPE=SECOND needs the property integrals (ref.perturbation)")
    mu, am = ref.perturbation("Mu"), ref.perturbation("L")
    omega = float(options["MY_OMEGA"])
    right = str(options["MYHAND"]) == "RIGHT"
    opts = cc._opts()
    opts.precision = ugx.PREC_FP64
    w = cc.out.write
    perts = {}
    try:
        w("\n\tSolving right hand perturbed CC amplitudes\n")
        for k, x in enumerate(CART):
            entry = "Mu" + x + _std_to_string(omega)
            entry_1 = "L" + x + _std_to_string(omega)
            w("\n\tCC RH Perturbed Wavefunction: %s\n" % entry)
            perts[entry] = cc.ctx.pert(mu[k], omega)
            _solve_pert(cc, perts[entry], 0, opts)
            w("\n\tCC RH Perturbed Wavefunction: %s\n" % entry_1)
            perts[entry_1] = cc.ctx.pert(am[k], omega)
            _solve_pert(cc, perts[entry_1], 0, opts)
            if omega != 0.0 and right:
                entry = "Mu" + x + _std_to_string(-omega)
                w("\n\tCC RH Perturbed Wavefunction: %s\n" % entry)
                perts[entry] = cc.ctx.pert(mu[k], -omega)
                _solve_pert(cc, perts[entry], 0, opts)
            if not right:
                w("\n\tSolving left hand perturbed CC amplitudes\n")
                w("\n\tCC LH Perturbed Wavefunction: %s\n" % entry)
                _solve_pert(cc, perts[entry], 1, opts)
                w("\n\tCC LH Perturbed Wavefunction: %s\n" % entry_1)
                _solve_pert(cc, perts[entry_1], 1, opts)
        polars, rots = {}, {}
        for p in range(3):
            for q in range(p + 1):
                gp = perts["Mu" + CART[p] + _std_to_string(omega)]
                gq = perts["Mu" + CART[q] + _std_to_string(omega)]
                if p == q:
                    polar = gp.linresp(gq)
                else:
                    polar = 0.5 * gp.linresp(gq) + 0.5 * gq.linresp(gp)
                polars["<<Mu_%s;Mu_%s>>" % (CART[p], CART[q])] = polar
                polars["<<Mu_%s;Mu_%s>>" % (CART[q], CART[p])] = polar
        for p in range(3):
            for q in range(3):
                gp = perts["Mu" + CART[p] + _std_to_string(omega)]
                gq = perts["L" + CART[q] + _std_to_string(omega)]
                rotation = 0.5 * gp.linresp(gq) - 0.5 * gq.linresp(gp)
                rots["<<Mu_%s;L_%s>>" % (CART[p], CART[q])] = -rotation
    finally:
        for g in perts.values():
            g.release()
    w("\n\tDipole Polarizabilities (au)\n\n")
    for label in sorted(polars):
        w("\t%s : %20.14f \n\n" % (label, polars[label]))
    w("\n\tOptical rotation (length gauge) (a.u)\n\n")
    for label in sorted(rots):
        w("\t%s : %20.14f \n\n" % (label, rots[label]))
    ref.polarizabilities = polars
    ref.rotations = rots


def ugacc(ref, options):
    """The plugin entry (ugacc.cc:75-241)."""
    if ref is None:
        raise PsiException("SCF has not been run yet!")
    if options.get("REFERENCE", "RHF") != "RHF" or getattr(ref, "reference", "RHF") != "RHF":
        raise PsiException("Only for use with RHF references.")
    if getattr(ref, "socc", 0):
        raise PsiException("UGACC is for closed-shell systems only.")
    dertype = options.get("DERTYPE", "NONE")
    # as in ugacc.cc:119-143, every DERTYPE other than NONE and FIRST runs
    # the linear response
    second = dertype not in ("NONE", "FIRST")
    if second and not hasattr(ref, "perturbation"):
        raise PsiException("DERTYPE=SECOND needs the property integrals (ref.perturbation)")
    if second and options.get("WFN") == "CCSD_T":
        # the reference's CCDensity copies the (T) density parts for every
        # CCSD_T run (ccdensity.cc:44-69), which CCWfn allocates only for
        # DERTYPE=FIRST (ccwfn.cc:108-118)
        raise PsiException("WFN=CCSD_T supports DERTYPE=NONE and FIRST only")
    cc = CCWfn(ref, options)
    try:
        ecc = cc.compute_energy()
        ref.cc_energy = ecc
        ref.cc_result = cc.result
        if dertype != "NONE":
            # HBAR, Lambda (with the (T) sources for CCSD_T at DERTYPE=FIRST:
            # cclambda.cc:181, :243) and the densities, ugacc.cc:118-141
            lo = cc._opts()
            lo.precision = ugx.PREC_FP64
            lr = cc.ctx.lambda_solve(lo)
            if not lr["converged"]:
                raise PsiException("Computation has not converged.")
            d = cc.ctx.density(which=())
            eone, etwo = d["e_one"], d["e_two"]
            eref = ref.escf
            wfn = "CCSD(T)" if cc.wfn == "CCSD_T" else cc.wfn
            w = cc.out.write
            w("\tOne-Electron Energy        = %20.14f\n" % eone)
            w("\tTwo-Electron Energy        = %20.14f\n" % etwo)
            w("\t%s Correlation Energy    = %20.14f (from density)\n" % (wfn, eone + etwo))
            w("\t%s Correlation Energy    = %20.14f (from ccwfn)\n" % (wfn, ecc))
            w("\t%s Total Energy          = %20.14f (from density)\n" % (wfn, eone + etwo + eref))
            w("\t%s Total Energy          = %20.14f (from ccwfn)\n" % (wfn, ecc + eref))
            ref.density_energy = eone + etwo
            ref.current_energy = ecc + eref
        if second:
            _linear_response(cc, ref, options)
    finally:
        cc.close()
    return ref
