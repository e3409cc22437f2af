"""Host restatement of the AMR solution transfer for the tests: Trixi's
refine!/coarsen! of the DG solution (src/callbacks_step/amr_dg2d.jl in
Trixi.jl) with the oracle's mortar operators, x first, then y, summed over
children; plan rows as TreeMesh.refine/coarsen give them."""
import numpy as np


def transfer(u, plan, F, R):
    """u [E_old, n, n, nv] (nodes [j, i]) -> [E_new, n, n, nv]; F = (forward_lower,
    forward_upper), R = (reverse_lower, reverse_upper), rows = output nodes."""
    out = np.zeros((len(plan),) + u.shape[1:])
    for e, (kind, a, b, c, d) in enumerate(plan):
        if kind == 0:
            out[e] = u[a]
        elif kind == 1:
            out[e] = np.einsum("jl,ik,lkv->jiv", F[b >> 1], F[b & 1], u[a])
        else:
            for q, src in enumerate((a, b, c, d)):
                out[e] += np.einsum("jl,ik,lkv->jiv", R[q >> 1], R[q & 1], u[src])
    return out
