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
ref.current_energy; DERTYPE=SECOND (the linear response) is refused.

``ref`` is any object with the attributes of the front end's
chem.scf.Reference: no, nv, fock (active Fock), blocks (six Dirac blocks),
escf, and optionally enuc, efzc, reference ("RHF"), socc (open-shell count).
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
            self.ctx.tgrad(which=())  # ccwfn.cc:252-255
        return etotal

    def close(self):
        self.ctx.close()


def ugacc(ref, options):
    """The plugin entry (ugacc.cc:75-117) for DERTYPE = NONE."""
    if ref is None:
        raise PsiException("SCF has not been run yet!")
    if options.get("REFERENCE", "RHF") != "RHF" or getattr(ref, "reference", "RHF") != "RHF":
        raise PsiException("Only for use with RHF references.")
    if getattr(ref, "socc", 0):
        raise PsiException("UGACC is for closed-shell systems only.")
    dertype = options.get("DERTYPE", "NONE")
    if dertype not in ("NONE", "FIRST"):
        raise PsiException("DERTYPE=%s needs the linear response (ccpert.cc, ccresp.cc), which this build "
                           "does not provide" % dertype)
    cc = CCWfn(ref, options)
    try:
        ecc = cc.compute_energy()
        ref.cc_energy = ecc
        ref.cc_result = cc.result
        if dertype == "FIRST":
            # HBAR, Lambda (with the (T) sources for CCSD_T) and the densities,
            # ugacc.cc:118-141
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
    finally:
        cc.close()
    return ref
