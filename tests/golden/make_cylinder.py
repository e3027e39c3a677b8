"""Writes the Bessel/Hankel-series golden fixture for the 2-D PEC cylinder.

Scene of the 2-D TMz cylinder workload (after test/cylinder_scatter_2d.cu:
28-183, in the 1000 x 1000, dl = 1 mm form): an infinite PEC cylinder of
radius 0.1 m along z in vacuum, lit by a plane wave E_z travelling along +x,
frequency 1 GHz.  Bistatic echo width (2-D radar cross section per unit
length) at the 360 angles linspace(-pi, pi, 360) of the xy-plane cut, phi
measured from the direction of travel (phi = 0 forward, pi backscatter);
the fixture also holds the cut for travel along phi_inc = pi/4 and 2.0 rad.

With exp(j w t) and E_z^i = exp(-j k x) = sum_n j^-n J_n(k rho) exp(j n phi),
the PEC boundary condition gives E_z^s = -sum_n j^-n a_n H2_n(k rho)
exp(j n phi), a_n = J_n(k a) / H2_n(k a); the large-argument form of H2_n
then gives (Balanis, "Advanced Engineering Electromagnetics", 2nd ed.,
sec. 11.5.1)

    sigma_2D(phi) = lim 2 pi rho |E^s|^2 / |E^i|^2
                  = (4 / k) | sum_{n=-N..N} a_n exp(j n phi) |^2,

with a_-n = a_n.  The numbers come from scipy.special, not from any FDTD
code.

python tests/golden/make_cylinder.py -> tests/golden/cylinder_pec_r0p1_1ghz.npz
"""
import os

import numpy as np
from scipy import special

C0 = 1.0 / np.sqrt(8.854187817e-12 * 4e-7 * np.pi)


def coefficients(ka, nmax):
    n = np.arange(0, nmax + 1)
    return special.jv(n, ka) / special.hankel2(n, ka)


def echo_width(radius, freq, phi):
    """sigma_2D(phi) in metres for a PEC cylinder, TMz (E along the axis)."""
    k = 2 * np.pi * freq / C0
    ka = k * radius
    nmax = int(np.ceil(ka + 4 * ka ** (1 / 3) + 2)) + 10
    a = coefficients(ka, nmax)
    s = a[0] + 2 * np.sum(a[1:, None] * np.cos(np.arange(1, nmax + 1)[:, None] * phi[None, :]), axis=0)
    return 4.0 / k * np.abs(s) ** 2


def main():
    phi = np.linspace(-np.pi, np.pi, 360)
    radius, freq = 0.1, 1e9
    sw = echo_width(radius, freq, phi)
    # the same cylinder lit along phi_inc = pi/4 (the reference driver's
    # TFSF2D angle) and 2.0 rad: the pattern turns with the incidence
    sw_45 = echo_width(radius, freq, phi - np.pi / 4)
    sw_2 = echo_width(radius, freq, phi - 2.0)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cylinder_pec_r0p1_1ghz.npz")
    np.savez(out, phi=phi, echo_width=sw, echo_width_phi_inc_0p785=sw_45, echo_width_phi_inc_2p0=sw_2,
             radius=radius, freq=freq)
    print(out, float(sw.max()), float(sw.min()))


if __name__ == "__main__":
    main()
