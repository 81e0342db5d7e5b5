"""Namelist variables and physical constants with the reference's names.

Mirrors namelist_n_constants.py:5-75 of the reference (the hot path's switches and
constants); the deep-learning entries (:77-84) live in lex_amd.sgs_net.  Values here are
the reference defaults; drivers overwrite them the way the reference's users edit the
namelist.
"""
# Solver setup (namelist_n_constants.py:6-14)
dt = 1.0
total_time = 1800
save_time = 300
sprint_n = 10
solver_opt = 2
n_sound = 32
alpha = 0.6

# Grid (namelist_n_constants.py:17-25)
dx = 100.0
dy = 100.0
dz = 100.0
nx = 240
ny = 240
nz = 120
ngx = 3
ngy = 3
ngz = 1

# Output files (namelist_n_constants.py:29-30)
file_name_format = "experiments/lex_out_%0.4i.nc"
base_file_name = "experiments/lex_reference_state.nc"

ic_option = 1
rand_opt = False
cor_opt = 0
fCor = 1.0e-4
sfc_opt = False
turb_opt = 0
damp_opt = False
rad_opt = False
pic_opt = False

# Bulk surface-flux scheme (namelist_n_constants.py:43-47; fixed in the device kernel)
atm_ocn_max_iter = 8
atm_ocn_min_wind = 0.25
sfc_z_ref = 10.0
sfc_z_t_ref = 2.0
Karman = 0.4

# Rayleigh damping layer (namelist_n_constants.py:54-55)
z_damping = 9600.0
rd_alpha = 1.0 / 120.0

# Smagorinsky constants (turbulence.py:432-433 hard-codes them; runtime values here)
cs = 0.18
prandtl = 1.0 / 3.0

# Physical constants (namelist_n_constants.py:64-75)
Rd = 287.04
Rv = 461.5
Cp = 1004.64
Cv = Cp - Rd
eps = Rd / Rv
Cpwv = 1810.0
Cpvir = Cpwv / Cp - 1.0
reps = Rv / Rd
repsm1 = Rv / Rd - 1.0
lat_vap = 2.501e6
g = 9.80616
p00 = 100000.0
