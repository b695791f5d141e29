// Setter/getter bodies shared by the lib_ipm and lib_irs C ABIs. Each ABI
// translation unit defines HANDLE (its opaque struct holding a MapMaker `m`)
// and DTYPE before including this file.
#pragma once

#include <exception>
#include <iostream>

#define MLENS_GUARD(expr, fail)                                   \
	try                                                           \
	{                                                             \
		expr;                                                     \
	}                                                             \
	catch (const std::exception& e)                               \
	{                                                             \
		std::cerr << "Error. " << e.what() << "\n";               \
		fail;                                                     \
	}                                                             \
	catch (const mlens::HipError& e)                              \
	{                                                             \
		std::cerr << "Error. " << e.msg << " (" << e.file << ":" << e.line << ")\n"; \
		fail;                                                     \
	}

#define MLENS_FIELD(name, type, member)                                   \
	void set_##name(HANDLE* self, type val) { self->m.member = val; }   \
	type get_##name(HANDLE* self) { return self->m.member; }

#define MLENS_STRING_FIELD(name, member)                                                       \
	void set_##name(HANDLE* self, const char* val) { self->m.member = val ? val : ""; }      \
	const char* get_##name(HANDLE* self) { return self->m.member.c_str(); }

#define MLENS_COMMON_FIELDS                                 \
	MLENS_FIELD(kappa_tot, DTYPE, kappa_tot)                \
	MLENS_FIELD(shear, DTYPE, shear)                        \
	MLENS_FIELD(kappa_star, DTYPE, kappa_star)              \
	MLENS_FIELD(theta_star, DTYPE, theta_star)              \
	MLENS_STRING_FIELD(mass_function, mass_function_str)    \
	MLENS_FIELD(m_solar, DTYPE, m_solar)                    \
	MLENS_FIELD(m_lower, DTYPE, m_lower)                    \
	MLENS_FIELD(m_upper, DTYPE, m_upper)                    \
	MLENS_FIELD(light_loss, DTYPE, light_loss)              \
	MLENS_FIELD(rectangular, int, rectangular)              \
	MLENS_FIELD(approx, int, approx)                        \
	MLENS_FIELD(safety_scale, DTYPE, safety_scale)          \
	MLENS_STRING_FIELD(starfile, starfile)                  \
	MLENS_FIELD(center_y1, DTYPE, center_y.re)              \
	MLENS_FIELD(center_y2, DTYPE, center_y.im)              \
	MLENS_FIELD(half_length_y1, DTYPE, half_length_y.re)    \
	MLENS_FIELD(half_length_y2, DTYPE, half_length_y.im)    \
	MLENS_FIELD(num_pixels_y1, int, num_pixels_y.re)        \
	MLENS_FIELD(num_pixels_y2, int, num_pixels_y.im)        \
	MLENS_FIELD(num_rays_y, int, num_rays_y)                \
	MLENS_FIELD(random_seed, int, random_seed)              \
	MLENS_FIELD(write_stars, int, write_stars)              \
	MLENS_FIELD(write_maps, int, write_maps)                \
	MLENS_FIELD(write_parities, int, write_parities)        \
	MLENS_FIELD(write_histograms, int, write_histograms)    \
	MLENS_STRING_FIELD(outfile_prefix, outfile_prefix)      \
	MLENS_FIELD(precision, int, precision)                  \
	MLENS_FIELD(m2l_mode, int, m2l_mode)                    \
	MLENS_FIELD(tile_rows, int, tile_rows)                  \
	void set_expansion_order(HANDLE* self, int val) { self->m.expansion_order_override = val; } \
	int get_expansion_order(HANDLE* self)                   \
	{                                                       \
		int p = self->m.get_expansion_order();              \
		return p > 0 ? p : self->m.expansion_order_override; \
	}

#define MLENS_COMMON_ENTRIES                                                                                   \
	double get_t_m2l(HANDLE* self) { return self->m.get_t_m2l(); }                                           \
	long long get_m2l_pairs(HANDLE* self) { return self->m.get_m2l_pairs(); }                                \
	int get_m2l_fp32(HANDLE* self) { return self->m.get_m2l_fp32(); }                                        \
	int get_num_stars(HANDLE* self) { return self->m.get_num_stars(); }                                      \
	DTYPE get_corner_x1(HANDLE* self) { return self->m.get_corner().re; }                                    \
	DTYPE get_corner_x2(HANDLE* self) { return self->m.get_corner().im; }                                    \
	DTYPE* get_stars(HANDLE* self) { return self->m.get_stars(); }                                           \
	int get_tree_levels(HANDLE* self) { return self->m.get_tree_levels(); }                                  \
	DTYPE get_alpha_error(HANDLE* self) { return self->m.get_alpha_error(); }                                \
	int get_num_ray_threads_1(HANDLE* self) { return self->m.get_num_ray_threads().re; }                     \
	int get_num_ray_threads_2(HANDLE* self) { return self->m.get_num_ray_threads().im; }                     \
	DTYPE get_root_half_length(HANDLE* self) { return self->m.get_root_half_length(); }                      \
	double get_mean_near_stars(HANDLE* self)                                                                   \
	{                                                                                                          \
		MLENS_GUARD(return self->m.mean_near_stars(), return 0)                                               \
	}                                                                                                          \
	long long get_cells_shot(HANDLE* self) { return self->m.get_cells_shot(); }                              \
	int get_taylor_smooth(HANDLE* self) { return self->m.get_taylor_smooth(); }                              \
	DTYPE get_center_x1(HANDLE* self) { return self->m.get_center_x().re; }                                  \
	DTYPE get_center_x2(HANDLE* self) { return self->m.get_center_x().im; }                                  \
	DTYPE get_half_length_x1(HANDLE* self) { return self->m.get_half_length_x().re; }                        \
	DTYPE get_half_length_x2(HANDLE* self) { return self->m.get_half_length_x().im; }                        \
	DTYPE get_ray_half_sep_1(HANDLE* self) { return self->m.get_ray_half_sep().re; }                         \
	DTYPE get_ray_half_sep_2(HANDLE* self) { return self->m.get_ray_half_sep().im; }                         \
	DTYPE get_kappa_star_actual(HANDLE* self) { return self->m.get_kappa_star_actual(); }                    \
	DTYPE get_mean_mass(HANDLE* self) { return self->m.get_mean_mass(); }                                    \
	DTYPE get_mean_mass2(HANDLE* self) { return self->m.get_mean_mass2(); }                                  \
	DTYPE get_mean_mass_actual(HANDLE* self) { return self->m.get_mean_mass_actual(); }                      \
	DTYPE get_mean_mass2_actual(HANDLE* self) { return self->m.get_mean_mass2_actual(); }                    \
	DTYPE get_mean_mass2_ln_mass(HANDLE* self) { return self->m.get_mean_mass2_ln_mass(); }                  \
	DTYPE get_mean_mass2_ln_mass_actual(HANDLE* self) { return self->m.get_mean_mass2_ln_mass_actual(); }    \
	DTYPE get_m_lower_actual(HANDLE* self) { return self->m.get_m_lower_actual(); }                          \
	DTYPE get_m_upper_actual(HANDLE* self) { return self->m.get_m_upper_actual(); }                          \
	int get_last_tile_rows(HANDLE* self) { return self->m.get_last_tile_rows(); }                            \
	DTYPE get_mu_ave(HANDLE* self) { return self->m.get_mu_ave(); }                                          \
	bool run(HANDLE* self, int verbose) { MLENS_GUARD(return self->m.run(verbose), return false) }           \
	bool save(HANDLE* self, int verbose) { MLENS_GUARD(return self->m.save(verbose), return false) }         \
	int get_num_nodes(HANDLE* self)                                                                            \
	{                                                                                                          \
		return self->m.engine() ? self->m.engine()->num_nodes() : 0;                                         \
	}
