"""Standard atomic weights (IUPAC 1999, Coplen, J. Phys. Chem. Ref. Data 30 (2001) 701-712), the
table the reference's equilibrium solver weighs elements with (Lib_CEA_data.f90:37-59); the solver
uses them for the element amounts b0 and the mixture enthalpy of its input
(Lib_CEA_setup.f90:126-139, 157-170)."""
import numpy as np

ATOMIC_WEIGHT = {
    "H": 1.00794, "D": 2.014102, "HE": 4.002602, "LI": 6.941, "BE": 9.012182, "B": 10.811, "C": 12.0107,
    "N": 14.0067, "O": 15.9994, "F": 18.9984032, "NE": 20.1797, "NA": 22.989770, "MG": 24.305,
    "AL": 26.981538, "SI": 28.0855, "P": 30.973761, "S": 32.065, "CL": 35.453, "AR": 39.948, "K": 39.0983,
    "CA": 40.078, "FE": 55.845, "XE": 131.293, "KR": 83.80,
}


def element_weights(elements):
    """Atomic weights for a list of element symbols (case-insensitive)."""
    try:
        return np.array([ATOMIC_WEIGHT[e.strip().upper()] for e in elements], dtype=np.float64)
    except KeyError as e:
        raise ValueError(f"no atomic weight for element {e}") from None
