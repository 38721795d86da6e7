-- 64^3 Sod shock tube on the unit cube (cgs, unit scales): rho = 1, p = 1
-- for x < 0.5 and rho = 0.125, p = 0.1 beyond; outflow on every face.
-- Keys are the reference's Lua parameter names (src/main.cpp:126-208).
Parameters = {
	Integration = {
		density_scale =   1.0,
		pressure_scale =  1.0,
		time_scale =      1.0,
		spatial_order =   1,        -- linear reconstruction (2nd order)
		temporal_order =  2,
		simulation_time = 1.0e3,
		radiation_on =    false,
		cooling_on =      false,
		ncheckpoints =    1,
	},
	Grid = {
		no_dimensions = 3,
		no_cells_x = 64,
		no_cells_y = 64,
		no_cells_z = 64,
		geometry = "cartesian",
		side_length = 1.0,
		left_boundary_condition_x = "outflow",
		left_boundary_condition_y = "outflow",
		left_boundary_condition_z = "outflow",
		right_boundary_condition_x = "outflow",
		right_boundary_condition_y = "outflow",
		right_boundary_condition_z = "outflow",
	},
	Hydrodynamics = {
		gamma = 1.67,
		density_floor = 1.0e-30,
		temperature_floor = 0,
		riemann_solver = "RotatedHLLC",
		slope_limiter = "albada",
	},
	Radiation = { mass_fraction_hydrogen = 1.0 },
}
