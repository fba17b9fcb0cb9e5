"""Physical constants and sentinel values used by the host layer.

Values follow xtrack/headers/constants.h:9-31, xtrack/particles/masses.py:12
and xtrack/particles/particles.py:24.
"""
C_LIGHT = 299792458.0
PI = 3.1415926535897932384626433832795028841971693993751
PROTON_MASS_EV = 938272088.16
ELECTRON_MASS_EV = 510998.95
LAST_INVALID_STATE = -999999999
# aperture default half-width (xtrack/beam_elements/_aperture_common.py:6)
UNLIMITED = 1e10
# global |x|, |y| limit of the default line config (xtrack/line.py:287)
GLOBAL_XY_LIMIT = 1.0
