"""riemann.static metadata (pyclaw/solver.py:230-238 reads num_eqn, num_waves, fwave)."""
num_eqn = {"acoustics_2D": 3, "euler_4wave_2D": 4, "shallow_roe_with_efix_2D": 3,
           "advection_2D": 1, "euler_3D": 5, "vc_advection_2D": 1,
           "vc_acoustics_2D": 3, "shallow_hlle_2D": 3, "acoustics_2D_ptwise": 3,
           "euler_5wave_2D": 5, "shallow_bathymetry_fwave_2D": 3}
num_waves = {"acoustics_2D": 2, "euler_4wave_2D": 4, "shallow_roe_with_efix_2D": 3,
             "advection_2D": 1, "euler_3D": 3, "vc_advection_vector_2D": 1,
             "vc_advection_2D": 1, "vc_acoustics_2D": 2, "shallow_hlle_2D": 2,
             "acoustics_2D_ptwise": 2, "euler_5wave_2D": 5, "shallow_bathymetry_fwave_2D": 3}
fwave = {k: False for k in num_waves}
fwave["shallow_bathymetry_fwave_2D"] = True
