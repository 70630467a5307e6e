"""The library's initial field of the dipole use case (ipic_field_init_host: initFields,
getDipoleFieldAt, initBcFields; field.h:43-138) on the host, against the oracle's
getDipoleFieldAt (or_dipole_field, pinned by the reference's dipole KATs field_test.cc:25-31,
301-312) at every node and against interpN2C's 8-node average (field.h:252-265) at every
centre.  The logged initial energies of the reference's runs are checked through the driver
(tests/test_driver_case.py)."""
import numpy as np
import pytest

import allscale_ipic3d_amd as G
import oracle as O

if not __import__("os").path.exists(G.LIB_PATH):
    pytest.skip("library not built", allow_module_level=True)


def _oracle_nodes(n, w, origin, drift, B0, center, radius, Bz):
    m = [k + 1 for k in n]
    E = np.zeros((*m, 3))
    B = np.zeros((*m, 3))
    e, bx, b = (np.zeros(3) for _ in range(3))
    for x in range(m[0]):
        for y in range(m[1]):
            for z in range(m[2]):
                loc = np.array([origin[0] + x * w[0], origin[1] + y * w[1], origin[2] + z * w[2]])
                O.lib().or_dipole_field(O._p(loc), O._p(O.f64(drift)), O._p(O.f64(B0)), O._p(O.f64(center)),
                                        float(radius), float(Bz), O._p(e), O._p(bx), O._p(b))
                E[x, y, z] = e
                B[x, y, z] = b
    return E, B


@pytest.mark.parametrize("n,w,origin,center,radius", [
    ((6, 5, 4), (0.5, 1.0, 2.0), (-1.5, -2.5, -4.0), (0.1, -0.2, 0.3), 0.9),
    ((4, 4, 4), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0), (2.0, 2.0, 2.0), 1.0),  # nodes on the planet surface
])
def test_host_field_init_matches_oracle_dipole(n, w, origin, center, radius):
    drift, B0, Bz = (0.02, -0.01, 0.005), (0.1, 0.0, 0.3), 2.5
    init = G.field_init(drift=drift, B0=B0, object_center=center, planet_radius=radius, external_Bz=Bz)
    got = G.field_init_host(G.desc_for(n, w, origin), init)
    E, B = _oracle_nodes(n, w, origin, drift, B0, center, radius, Bz)
    np.testing.assert_allclose(got["E"], E[:n[0], :n[1], :n[2]].reshape(-1, 3), rtol=1e-15, atol=0)
    np.testing.assert_allclose(got["Bn"], B[:n[0], :n[1], :n[2]].reshape(-1, 3), rtol=1e-14, atol=1e-300)
    Bc = np.zeros((*n, 3))
    for a in (0, 1):
        for b in (0, 1):
            for c in (0, 1):
                Bc += B[a:a + n[0], b:b + n[1], c:c + n[2]]
    np.testing.assert_allclose(got["Bc"], 0.125 * Bc.reshape(-1, 3), rtol=1e-14, atol=1e-300)
    # E = -drift x B0 everywhere (field.h:47, 52)
    np.testing.assert_allclose(got["E"][0], -np.cross(drift, B0), rtol=1e-15)
    # energies: getFieldEnergy over the (N+1)^3 nodes, of E and of B + Bext = 2 B - B0
    vol = 0.5 * w[0] * w[1] * w[2] / (16.0 * np.arctan(1.0))
    eE = vol * np.sum(E * E)
    Bext = B - np.asarray(B0)
    eB = vol * np.sum((B + Bext) ** 2)
    assert got["energies"][0] == pytest.approx(eE, rel=1e-13)
    assert got["energies"][1] == pytest.approx(eB, rel=1e-13)


def test_dipole_moment_convention():
    """dipole_field_init(moment, ...) gives B = moment (3xz, 3yz, 2z^2 - x^2 - y^2) / r^5 at a
    node outside the radius and 0 inside."""
    n = (8, 8, 8)
    init = G.dipole_field_init(1.0, (4.0, 4.0, 4.0), 2.0)
    got = G.field_init_host(G.desc_for(n), init)
    Bn = got["Bn"].reshape(8, 8, 8, 3)
    d = np.array([7.0, 6.0, 1.0]) - 4.0
    r2 = d @ d
    ref = np.array([3 * d[0] * d[2], 3 * d[1] * d[2], 2 * d[2] ** 2 - d[0] ** 2 - d[1] ** 2]) / r2 ** 2.5
    np.testing.assert_allclose(Bn[7, 6, 1], ref, rtol=1e-14)
    assert np.all(Bn[4, 4, 4] == 0.0) and np.all(Bn[5, 4, 4] == 0.0)  # r = 0 and r = 1 < 2


def test_host_field_init_refuses_bad_geometry():
    d = G.desc_for((4, 4, 4))
    d.cell_width[1] = 0.0
    with pytest.raises(G.IpicError, match="cell width"):
        G.field_init_host(d, G.field_init())
    with pytest.raises(G.IpicError, match="planet radius"):
        G.field_init_host(G.desc_for(4), G.field_init(planet_radius=-1.0))
