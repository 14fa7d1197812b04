"""Input fingerprint of a synthetic workload, stored beside each golden vector file."""
import hashlib

import numpy as np


def input_hash(w) -> str:
    h = hashlib.sha256()
    for a in (w.basis.kinetic, w.basis.cart, w.local_r, w.atom_pos, w.psi):
        h.update(np.ascontiguousarray(a).tobytes())
    for s in range(len(w.species)):
        h.update(np.ascontiguousarray(w.phi(s)).tobytes())
    return h.hexdigest()
