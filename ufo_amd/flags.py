"""Simulation flags.

Same bit values and names as the reference (ufo/methods.py:8-14), so user
code that combines them (``KICK | DOUBLE_PRECISION | FIVED``) keeps working.
"""

LINEAR = 1 << 0            # keep only multipole orders <= 1 (ufo/methods.py:50,64)
FIVED = 1 << 1             # momentum fixed to the ``dp`` argument, z = 0
EXACT = 1 << 2             # exact-Hamiltonian drift (ufo/methods.py:29-37)
KICK = 1 << 3              # teapot slicing of thick magnets (ufo/methods.py:79-97)
RADIATION = 1 << 4         # accepted and ignored, as in the reference
DOUBLE_PRECISION = 1 << 5  # FP64 arithmetic (the default is FP32)
ACHROMATIC = 1 << 6        # drop the 1/(1+dp) chromatic factor of kicks and bends

ALL_FLAGS = LINEAR | FIVED | EXACT | KICK | RADIATION | DOUBLE_PRECISION | ACHROMATIC
