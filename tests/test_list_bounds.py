"""The two bounds the contact test relies on, checked on the oracle's own
trajectory (CPU only), for the scene of the aging-lists parity test
(`parity.aging_lists_scene`: sparse spheres at one speed in random
directions, skin 0.3 d_max, periodic in x, y and z).

* Candidate lists built with skin s stay complete until the box's summed
  largest per-step moves since the build exceed s/2 (every pair that
  touches was listed).
* A particle whose smallest listed gap |xi - xj| - (ri + rj) at the build
  exceeds twice those summed moves touches nothing (the contact test skips
  it); with once the summed moves the scene does have particles that touch,
  so it tells the two apart.

The rebuild schedule and the lists are recomputed here with a k-d tree from
the oracle's states; nothing of the HIP library is used."""
import numpy as np
from scipy.spatial import cKDTree

import oracle
from parity import aging_lists_scene
from sdem_amd import native


def gaps(x, r, L, skin):
    """Smallest gap to a candidate (pairs within ri + rj + skin) per
    particle, +inf for particles without candidates."""
    t = cKDTree(x, boxsize=L)
    pairs = t.query_pairs(2 * r.max() + skin, output_type="ndarray")
    i, j = pairs[:, 0], pairs[:, 1]
    d = x[i] - x[j]
    d -= L * np.round(d / L)
    g = np.sqrt((d * d).sum(1)) - r[i] - r[j]
    keep = g <= skin
    out = np.full(len(r), np.inf)
    np.minimum.at(out, i[keep], g[keep])
    np.minimum.at(out, j[keep], g[keep])
    return out


def test_list_bounds_on_the_oracle_trajectory(oracle_lib):
    sc = aging_lists_scene()
    o = sc["options"]
    lo, L = np.array(o["domain_min"]), np.array(o["domain_max"]) - np.array(o["domain_min"])
    ref = native.build_engine(oracle_lib, sc, prefix=oracle.PREFIX)
    st = ref.get_state()
    r = st["posr"][:, 3]
    skin = 0.3 * 2 * r.max()
    wrap = lambda p: np.mod(p - lo, L)  # noqa: E731
    x = wrap(st["posr"][:, :3])
    g = gaps(x, r, L, skin)
    disp, builds = 0.0, 1
    missed1 = touching = 0
    for step in range(80):
        if disp > 0.5 * skin:  # the lists may be stale: build them again
            g, disp, builds = gaps(x, r, L, skin), 0.0, builds + 1
        ref.iterate(1)
        ct = ref.get_contacts()
        pp = ct["kind"] != native.KIND_PW
        ids = st["id"]
        hit = np.isin(ids, np.concatenate([ct["owner"][pp], ct["partner"][pp]]))
        touching += int(hit.sum())
        # contacts of this step were found on x (the state before it)
        assert np.all(np.isfinite(g[hit])), "a touching pair was not listed"
        assert np.all(g[hit] <= 2.0 * disp), "the gap bound would skip a touching particle"
        missed1 += int(np.sum(g[hit] > disp))
        st = ref.get_state()
        assert np.array_equal(st["id"], ids)
        xn = wrap(st["posr"][:, :3])
        d = xn - x
        d -= L * np.round(d / L)
        disp += float(np.sqrt((d * d).sum(1)).max())
        x = xn
    assert touching > 20 and builds > 2, (touching, builds)
    assert missed1 > 0, "the scene does not tell twice the summed moves from once"
