"""Physical constants and search defaults of the TLS search engine.

The numerical values are those of the reference (src/gputls/constants.py:33-171 and the
`#define` block of src/gputls/GPUFun.cu:1-21): every grid, template and window bound
derived from them has to reproduce the reference's values, so the numbers are kept as-is.
"""

# astrophysical constants (src/gputls/constants.py:33-38)
G = 6.673e-11            # m^3 kg^-1 s^-2
R_SUN = 695508000        # m
R_EARTH = 6371000        # m
R_JUP = 69911000         # m
M_SUN = 1.989 * 10 ** 30  # kg
SECONDS_PER_DAY = 86400

# search defaults (src/gputls/constants.py:41-55)
TRANSIT_DEPTH_MIN = 10 * 10 ** -6
NUMERICAL_STABILITY_CUTOFF = 0.01 * 10 ** -6
R_STAR = 1.0
M_STAR = 1.0
OVERSAMPLING_FACTOR = 3
N_TRANSITS_MIN = 2
M_STAR_MIN = 0.1
M_STAR_MAX = 1.0
R_STAR_MIN = 0.13
R_STAR_MAX = 3.5
DURATION_GRID_STEP = 1.1

# default template: quadratic limb darkening of a G2V star (src/gputls/constants.py:60-67)
DEFAULT_U = (0.4804, 0.1867)
DEFAULT_LIMB_DARK = "quadratic"
DEFAULT_ECC = 0
DEFAULT_W = 90
DEFAULT_PERIOD = 12.9
DEFAULT_RP = 0.03
DEFAULT_A = 23.1
DEFAULT_INC = 89.21

# grazing and box templates (src/gputls/constants.py:69-79)
GRAZING_B = 0.99
BOX_PERIOD = 29
BOX_RP = 0.1
BOX_A = 26.9
BOX_B = 0
BOX_INC = 90
BOX_U = (0,)
BOX_LIMB_DARK = "linear"

SIGNAL_DEPTH = 0.5                      # src/gputls/constants.py:84
FRACTIONAL_TRANSIT_DURATION_MAX = 0.12  # src/gputls/constants.py:91
SUPERSAMPLE_SIZE = 10000                # src/gputls/constants.py:102
OVERSAMPLE_MODEL_LIGHT_CURVE = 5        # src/gputls/constants.py:103
SDE_MEDIAN_KERNEL_SIZE = 30             # src/gputls/constants.py:113
T0_FIT_MARGIN = 0.125                   # src/gputls/constants.py:122
MINIMUM_PERIOD_GRID_SIZE = 100          # src/gputls/constants.py:133

# Per-period width window of the search (src/gputls/GPUFun.cu:5-21). These differ from the
# duration-grid constants above on purpose: the reference uses a wider stellar range here.
WIN_R_STAR_MIN = 0.05
WIN_R_STAR_MAX = 4.0
WIN_FRACTIONAL_MAX = 0.15
WIN_PI_GM_MAX = 416970
WIN_PI_GM_MIN = 20848
WIN_SCALE = 1000000000000000.0

# value of the skip factor when T0_fit_margin == 0 (src/gputls/core.py:598-599: 0x7f800000)
SKIP_NONE = 0x7F800000

# keyword arguments accepted by power() (src/gputls/constants.py:138-171), plus the
# engine's own `mode` and `device` selectors
VALID_PARAMETERS = (
    "R_star", "R_star_min", "R_star_max", "M_star", "M_star_min", "M_star_max",
    "period_min", "period_max", "periods", "n_transits_min", "GPUDeviceID",
    "GPUDeviceIDs", "legacy", "fast", "per", "rp", "a", "inc", "b", "ecc", "w", "u",
    "limb_dark", "duration_grid_step", "transit_depth_min", "oversampling_factor",
    "T0_fit_margin", "use_threads", "show_progress_bar", "transit_template", "verbose",
    "bar_location", "useLocalPTXCUBIN", "mode",
)

VERSION = "0.1.0"
TLS_VERSION = "GTLS-AMD " + VERSION + " (MI355X-native TLS search)"
