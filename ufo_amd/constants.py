"""Physical constants (CODATA values; the reference keeps the same names in
ufo/constants.py:9-14)."""

speed_of_light = 299792458.0              # m/s
electron_mass = 0.51099895 * 1e6          # eV; this product (510998.94999999995) is the reference's value
proton_mass = 938.27208816e6              # eV
elementary_charge = 1.602176634e-19       # C
vacuum_permittivity = 8.8541878128e-12    # F/m
classical_electron_radius = 2.8179403262e-15  # m

# names and values of ufo/constants.py:11-14 (the reference's rounded
# radiation constants, kept so emittance and energy loss agree with it)
electron_radius = 2.81794092e-15          # m
quantum_radiation_constant = 3.832e-13    # m, C_q = 55 hbar c / (32 sqrt(3) m c^2)
radiation_loss_constant = 14.085e-6       # m / GeV^3, C_gamma / (2 pi)
