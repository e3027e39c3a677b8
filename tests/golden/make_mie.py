"""Writes the Mie-series golden fixture for the dielectric-sphere driver.

Sphere of the reference driver (test/dielectric_sphere_scatter.cu:34-38):
radius 0.1 m, eps_r = 3, mu_r = 2, in vacuum, plane wave polarised along x
travelling along +z, frequency 1 GHz.  Bistatic RCS in the xz plane
(E-plane, sigma_theta = 4 pi |S2|^2 / k^2) and the yz plane (H-plane,
sigma_phi = 4 pi |S1|^2 / k^2) at the 360 angles linspace(-pi, pi, 360) the
driver asks processFarField for.  Coefficients: Bohren & Huffman,
"Absorption and Scattering of Light by Small Particles" (1983), eq. 4.53
with a magnetic sphere (mu_1 = mu_r, medium mu = 1); angular functions
pi_n, tau_n by their upward recurrences (eq. 4.47).  The numbers come from
scipy.special, not from any FDTD code.

python tests/golden/make_mie.py  ->  tests/golden/mie_sphere_eps3_mu2_r0p1_1ghz.npz
"""
import os

import numpy as np
from scipy import special

C0 = 1.0 / np.sqrt(8.854187817e-12 * 4e-7 * np.pi)


def mie_coefficients(x, eps_r, mu_r, nmax):
    m = np.sqrt(eps_r * mu_r)
    mx = m * x
    n = np.arange(1, nmax + 1)
    jn_x = special.spherical_jn(n, x)
    jn_x_d = special.spherical_jn(n, x, derivative=True)
    yn_x = special.spherical_yn(n, x)
    yn_x_d = special.spherical_yn(n, x, derivative=True)
    jn_mx = special.spherical_jn(n, mx)
    jn_mx_d = special.spherical_jn(n, mx, derivative=True)
    hn_x = jn_x + 1j * yn_x
    hn_x_d = jn_x_d + 1j * yn_x_d
    # [z f(z)]' = f(z) + z f'(z)
    xj_d = jn_x + x * jn_x_d
    xh_d = hn_x + x * hn_x_d
    mxj_d = jn_mx + mx * jn_mx_d
    mu, mu1 = 1.0, mu_r
    a = (mu * m**2 * jn_mx * xj_d - mu1 * jn_x * mxj_d) / (mu * m**2 * jn_mx * xh_d - mu1 * hn_x * mxj_d)
    b = (mu1 * jn_mx * xj_d - mu * jn_x * mxj_d) / (mu1 * jn_mx * xh_d - mu * hn_x * mxj_d)
    return a, b


def amplitudes(theta, a, b):
    mu = np.cos(theta)
    nmax = len(a)
    pi_prev = np.zeros_like(mu)
    pi_cur = np.ones_like(mu)
    s1 = np.zeros(theta.shape, complex)
    s2 = np.zeros(theta.shape, complex)
    for n in range(1, nmax + 1):
        tau = n * mu * pi_cur - (n + 1) * pi_prev
        f = (2 * n + 1) / (n * (n + 1))
        s1 += f * (a[n - 1] * pi_cur + b[n - 1] * tau)
        s2 += f * (a[n - 1] * tau + b[n - 1] * pi_cur)
        pi_next = ((2 * n + 1) * mu * pi_cur - (n + 1) * pi_prev) / n
        pi_prev, pi_cur = pi_cur, pi_next
    return s1, s2


def rcs(radius, eps_r, mu_r, freq, theta):
    k = 2 * np.pi * freq / C0
    x = k * radius
    nmax = int(np.ceil(x + 4 * x ** (1 / 3) + 2)) + 10
    a, b = mie_coefficients(x, eps_r, mu_r, nmax)
    s1, s2 = amplitudes(np.abs(theta), a, b)
    return 4 * np.pi * np.abs(s2) ** 2 / k**2, 4 * np.pi * np.abs(s1) ** 2 / k**2


def main():
    theta = np.linspace(-np.pi, np.pi, 360)
    xz, yz = rcs(0.1, 3.0, 2.0, 1e9, theta)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                       "mie_sphere_eps3_mu2_r0p1_1ghz.npz")
    np.savez(out, theta=theta, rcs_xz=xz, rcs_yz=yz)
    print(out, float(xz.max()), float(yz.max()))


if __name__ == "__main__":
    main()
