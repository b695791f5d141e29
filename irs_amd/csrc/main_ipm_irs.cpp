// Command-line programs: the options, defaults, messages and exit codes of
// the reference's executables (src/main_ipm_irs.cu:41-606, src/main_ccf.cu:29-523,
// src/main_ncc.cu:24-298), built as irs_amd/bin/ipm (float), ipm_double and irs
// (float) over the map engine, irs_amd/bin/ccf (double) over the critical-curve
// solver and irs_amd/bin/ncc (double) over the caustic-crossing counter.
//
// Every option takes a value: `ipm -k 0.45 -y 0.5 -s 0.2 -mf kroupa ...`.
// Star-field options are ignored when a star file is given; -s is ignored
// when -ks is given and otherwise sets kappa_star = (1 - s) kappa_tot after
// all options are read.
#if defined(MLENS_CLI_CCF)
#include "ccf_maker.hpp"
#elif defined(MLENS_CLI_NCC)
#include "ncc_maker.hpp"
#else
#include "map_maker.hpp"
#endif

#include <algorithm>
#include <cctype>
#include <functional>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#if !defined(MLENS_CLI_IPM) && !defined(MLENS_CLI_IRS) && !defined(MLENS_CLI_CCF) && !defined(MLENS_CLI_NCC)
#error "define MLENS_CLI_IPM, MLENS_CLI_IRS, MLENS_CLI_CCF or MLENS_CLI_NCC"
#endif
#if defined(MLENS_CLI_NCC)
#define MLENS_CLI_STAR_FIELD 0
#else
#define MLENS_CLI_STAR_FIELD 1
#endif
#ifndef MLENS_CLI_DTYPE
#define MLENS_CLI_DTYPE float
#endif

using dtype = MLENS_CLI_DTYPE;
#if defined(MLENS_CLI_CCF)
using Maker = mlens::CCFMaker<dtype>;
#elif defined(MLENS_CLI_NCC)
using Maker = mlens::NCCMaker<dtype>;
#elif defined(MLENS_CLI_IPM)
using Maker = mlens::MapMaker<dtype, true>;
#else
using Maker = mlens::MapMaker<dtype, false>;
#endif

namespace
{

Maker maker;
int verbose = 1;
#if MLENS_CLI_STAR_FIELD
dtype smooth_fraction = static_cast<dtype>(1 - maker.kappa_star / maker.kappa_tot);
#endif

template <typename V>
std::string show(const V& v)
{
	std::ostringstream s;
	s << v;
	return s.str();
}
template <typename V>
std::string show(const mlens::Complex<V>& v)
{
	std::ostringstream s;
	s << "(" << v.re << ", " << v.im << ")";
	return s.str();
}

#if MLENS_CLI_STAR_FIELD
std::string lowercase(std::string s)
{
	std::transform(s.begin(), s.end(), s.begin(), [](unsigned char c) { return char(std::tolower(c)); });
	return s;
}
#endif

enum Kind
{
	REAL,
	INTEGER,
	TEXT
};

struct Option
{
	const char* shrt;
	const char* lng;
	const char* name;                           // parameter name in messages
	Kind kind;
	bool star_field;                            // ignored when a star file is given
	std::function<void(double, int, const std::string&)> set;
	std::function<std::string()> get;
};

template <typename F>
Option opt(const char* s, const char* l, const char* name, Kind k, bool star, F&& set, std::function<std::string()> get)
{
	return Option{s, l, name, k, star, std::forward<F>(set), std::move(get)};
}

#if defined(MLENS_CLI_NCC)
// main_ncc.cu:24-39
std::vector<Option> options()
{
	Maker& m = maker;
	return {
		opt("-ip", "--infile_prefix", "infile_prefix", TEXT, false,
			[&](double, int, const std::string& s) { m.infile_prefix = s; }, [&] { return m.infile_prefix; }),
		opt("-cy1", "--center_y1", "center_y1", REAL, false, [&](double d, int, const std::string&) { m.center_y.re = dtype(d); },
			[&] { return show(m.center_y.re); }),
		opt("-cy2", "--center_y2", "center_y2", REAL, false, [&](double d, int, const std::string&) { m.center_y.im = dtype(d); },
			[&] { return show(m.center_y.im); }),
		opt("-hly1", "--half_length_y1", "half_length_y1", REAL, false,
			[&](double d, int, const std::string&) { m.half_length_y.re = dtype(d); }, [&] { return show(m.half_length_y.re); }),
		opt("-hly2", "--half_length_y2", "half_length_y2", REAL, false,
			[&](double d, int, const std::string&) { m.half_length_y.im = dtype(d); }, [&] { return show(m.half_length_y.im); }),
		opt("-npy1", "--num_pixels_y1", "num_pixels_y1", INTEGER, false,
			[&](double, int i, const std::string&) { m.num_pixels_y.re = i; }, [&] { return show(m.num_pixels_y.re); }),
		opt("-npy2", "--num_pixels_y2", "num_pixels_y2", INTEGER, false,
			[&](double, int i, const std::string&) { m.num_pixels_y.im = i; }, [&] { return show(m.num_pixels_y.im); }),
		opt("-os", "--over_sample", "over_sample", INTEGER, false, [&](double, int i, const std::string&) { m.over_sample = i; },
			[&] { return show(m.over_sample); }),
		opt("-wm", "--write_maps", "write_maps", INTEGER, false, [&](double, int i, const std::string&) { m.write_maps = i; },
			[&] { return show(m.write_maps); }),
		opt("-wh", "--write_histograms", "write_histograms", INTEGER, false,
			[&](double, int i, const std::string&) { m.write_histograms = i; }, [&] { return show(m.write_histograms); }),
		opt("-o", "--outfile_prefix", "outfile_prefix", TEXT, false,
			[&](double, int, const std::string& s) { m.outfile_prefix = s; }, [&] { return m.outfile_prefix; }),
	};
}
#elif defined(MLENS_CLI_CCF)
// main_ccf.cu:29-55
std::vector<Option> options()
{
	Maker& m = maker;
	return {
		opt("-k", "--kappa_tot", "kappa_tot", REAL, false, [&](double d, int, const std::string&) { m.kappa_tot = dtype(d); },
			[&] { return show(m.kappa_tot); }),
		opt("-y", "--shear", "shear", REAL, false, [&](double d, int, const std::string&) { m.shear = dtype(d); },
			[&] { return show(m.shear); }),
		opt("-s", "--smooth_fraction", "smooth_fraction", REAL, true,
			[&](double d, int, const std::string&) { smooth_fraction = dtype(d); }, [&] { return show(smooth_fraction); }),
		opt("-ks", "--kappa_star", "kappa_star", REAL, true, [&](double d, int, const std::string&) { m.kappa_star = dtype(d); },
			[&] { return show(m.kappa_star); }),
		opt("-t", "--theta_star", "theta_star", REAL, true, [&](double d, int, const std::string&) { m.theta_star = dtype(d); },
			[&] { return show(m.theta_star); }),
		opt("-mf", "--mass_function", "mass_function", TEXT, true,
			[&](double, int, const std::string& s) { m.mass_function_str = lowercase(s); }, [&] { return m.mass_function_str; }),
		opt("-ms", "--m_solar", "m_solar", REAL, false, [&](double d, int, const std::string&) { m.m_solar = dtype(d); },
			[&] { return show(m.m_solar); }),
		opt("-ml", "--m_lower", "m_lower", REAL, true, [&](double d, int, const std::string&) { m.m_lower = dtype(d); },
			[&] { return show(m.m_lower); }),
		opt("-mh", "--m_upper", "m_upper", REAL, true, [&](double d, int, const std::string&) { m.m_upper = dtype(d); },
			[&] { return show(m.m_upper); }),
		opt("-r", "--rectangular", "rectangular", INTEGER, true, [&](double, int i, const std::string&) { m.rectangular = i; },
			[&] { return show(m.rectangular); }),
		opt("-a", "--approx", "approx", INTEGER, false, [&](double, int i, const std::string&) { m.approx = i; },
			[&] { return show(m.approx); }),
		opt("-ss", "--safety_scale", "safety_scale", REAL, false,
			[&](double d, int, const std::string&) { m.safety_scale = dtype(d); }, [&] { return show(m.safety_scale); }),
		opt("-ns", "--num_stars", "num_stars", INTEGER, true, [&](double, int i, const std::string&) { m.num_stars = i; },
			[&] { return show(m.num_stars); }),
		opt("-sf", "--starfile", "starfile", TEXT, false, [&](double, int, const std::string& s) { m.starfile = s; },
			[&] { return m.starfile; }),
		opt("-np", "--num_phi", "num_phi", INTEGER, false, [&](double, int i, const std::string&) { m.num_phi = i; },
			[&] { return show(m.num_phi); }),
		opt("-nb", "--num_branches", "num_branches", INTEGER, false,
			[&](double, int i, const std::string&) { m.num_branches = i; }, [&] { return show(m.num_branches); }),
		opt("-rs", "--random_seed", "random_seed", INTEGER, true, [&](double, int i, const std::string&) { m.random_seed = i; },
			[&] { return show(m.random_seed); }),
		opt("-ws", "--write_stars", "write_stars", INTEGER, false, [&](double, int i, const std::string&) { m.write_stars = i; },
			[&] { return show(m.write_stars); }),
		opt("-wcc", "--write_critical_curves", "write_critical_curves", INTEGER, false,
			[&](double, int i, const std::string&) { m.write_critical_curves = i; },
			[&] { return show(m.write_critical_curves); }),
		opt("-wc", "--write_caustics", "write_caustics", INTEGER, false,
			[&](double, int i, const std::string&) { m.write_caustics = i; }, [&] { return show(m.write_caustics); }),
		opt("-wls", "--write_length_scales", "write_length_scales", INTEGER, false,
			[&](double, int i, const std::string&) { m.write_mu_length_scales = i; },
			[&] { return show(m.write_mu_length_scales); }),
		opt("-o", "--outfile_prefix", "outfile_prefix", TEXT, false,
			[&](double, int, const std::string& s) { m.outfile_prefix = s; }, [&] { return m.outfile_prefix; }),
	};
}
#else
std::vector<Option> options()
{
	Maker& m = maker;
	return {
		opt("-k", "--kappa_tot", "kappa_tot", REAL, false, [&](double d, int, const std::string&) { m.kappa_tot = dtype(d); },
			[&] { return show(m.kappa_tot); }),
		opt("-y", "--shear", "shear", REAL, false, [&](double d, int, const std::string&) { m.shear = dtype(d); },
			[&] { return show(m.shear); }),
		opt("-s", "--smooth_fraction", "smooth_fraction", REAL, true,
			[&](double d, int, const std::string&) { smooth_fraction = dtype(d); }, [&] { return show(smooth_fraction); }),
		opt("-ks", "--kappa_star", "kappa_star", REAL, true, [&](double d, int, const std::string&) { m.kappa_star = dtype(d); },
			[&] { return show(m.kappa_star); }),
		opt("-t", "--theta_star", "theta_star", REAL, true, [&](double d, int, const std::string&) { m.theta_star = dtype(d); },
			[&] { return show(m.theta_star); }),
		opt("-mf", "--mass_function", "mass_function", TEXT, true,
			[&](double, int, const std::string& s) { m.mass_function_str = lowercase(s); }, [&] { return m.mass_function_str; }),
		opt("-ms", "--m_solar", "m_solar", REAL, true, [&](double d, int, const std::string&) { m.m_solar = dtype(d); },
			[&] { return show(m.m_solar); }),
		opt("-ml", "--m_lower", "m_lower", REAL, true, [&](double d, int, const std::string&) { m.m_lower = dtype(d); },
			[&] { return show(m.m_lower); }),
		opt("-mh", "--m_upper", "m_upper", REAL, true, [&](double d, int, const std::string&) { m.m_upper = dtype(d); },
			[&] { return show(m.m_upper); }),
		opt("-ll", "--light_loss", "light_loss", REAL, false, [&](double d, int, const std::string&) { m.light_loss = dtype(d); },
			[&] { return show(m.light_loss); }),
		opt("-r", "--rectangular", "rectangular", INTEGER, true, [&](double, int i, const std::string&) { m.rectangular = i; },
			[&] { return show(m.rectangular); }),
		opt("-a", "--approx", "approx", INTEGER, false, [&](double, int i, const std::string&) { m.approx = i; },
			[&] { return show(m.approx); }),
		opt("-ss", "--safety_scale", "safety_scale", REAL, false,
			[&](double d, int, const std::string&) { m.safety_scale = dtype(d); }, [&] { return show(m.safety_scale); }),
		opt("-sf", "--starfile", "starfile", TEXT, false, [&](double, int, const std::string& s) { m.starfile = s; },
			[&] { return m.starfile; }),
		opt("-cy1", "--center_y1", "center_y1", REAL, false, [&](double d, int, const std::string&) { m.center_y.re = dtype(d); },
			[&] { return show(m.center_y.re); }),
		opt("-cy2", "--center_y2", "center_y2", REAL, false, [&](double d, int, const std::string&) { m.center_y.im = dtype(d); },
			[&] { return show(m.center_y.im); }),
		opt("-hly1", "--half_length_y1", "half_length_y1", REAL, false,
			[&](double d, int, const std::string&) { m.half_length_y.re = dtype(d); }, [&] { return show(m.half_length_y.re); }),
		opt("-hly2", "--half_length_y2", "half_length_y2", REAL, false,
			[&](double d, int, const std::string&) { m.half_length_y.im = dtype(d); }, [&] { return show(m.half_length_y.im); }),
		opt("-npy1", "--num_pixels_y1", "num_pixels_y1", INTEGER, false,
			[&](double, int i, const std::string&) { m.num_pixels_y.re = i; }, [&] { return show(m.num_pixels_y.re); }),
		opt("-npy2", "--num_pixels_y2", "num_pixels_y2", INTEGER, false,
			[&](double, int i, const std::string&) { m.num_pixels_y.im = i; }, [&] { return show(m.num_pixels_y.im); }),
		opt("-nry", "--num_rays_y", "num_rays_y", INTEGER, false, [&](double, int i, const std::string&) { m.num_rays_y = i; },
			[&] { return show(m.num_rays_y); }),
		opt("-rs", "--random_seed", "random_seed", INTEGER, true, [&](double, int i, const std::string&) { m.random_seed = i; },
			[&] { return show(m.random_seed); }),
		opt("-ws", "--write_stars", "write_stars", INTEGER, false, [&](double, int i, const std::string&) { m.write_stars = i; },
			[&] { return show(m.write_stars); }),
		opt("-wm", "--write_maps", "write_maps", INTEGER, false, [&](double, int i, const std::string&) { m.write_maps = i; },
			[&] { return show(m.write_maps); }),
		opt("-wp", "--write_parities", "write_parities", INTEGER, false,
			[&](double, int i, const std::string&) { m.write_parities = i; }, [&] { return show(m.write_parities); }),
		opt("-wh", "--write_histograms", "write_histograms", INTEGER, false,
			[&](double, int i, const std::string&) { m.write_histograms = i; }, [&] { return show(m.write_histograms); }),
		opt("-o", "--outfile_prefix", "outfile_prefix", TEXT, false,
			[&](double, int, const std::string& s) { m.outfile_prefix = s; }, [&] { return m.outfile_prefix; }),
		// extensions: long options only
		opt("", "--precision", "precision", INTEGER, false, [&](double, int i, const std::string&) { m.precision = i; },
			[&] { return show(m.precision); }),
		opt("", "--expansion_order", "expansion_order", INTEGER, false,
			[&](double, int i, const std::string&) { m.expansion_order_override = i; },
			[&] { return show(m.expansion_order_override); }),
		opt("", "--m2l_mode", "m2l_mode", INTEGER, false, [&](double, int i, const std::string&) { m.m2l_mode = i; },
			[&] { return show(m.m2l_mode); }),
	};
}
#endif

void usage(const char* name)
{
	std::cout << "Usage: " << (name ? name : "programname") << " opt1 val1 opt2 val2 opt3 val3 ...\n\nOptions:\n";
	std::cout << "  -h,--help                Show this help message.\n";
	std::cout << "  -v,--verbose             Verbosity from 0 (none) to 1, 2 or 3. Default value: " << verbose << "\n";
	for (const Option& o : options())
	{
		std::string flags = o.shrt[0] ? std::string(o.shrt) + "," + o.lng : std::string("   ") + o.lng;
		flags.resize(std::max<size_t>(flags.size() + 1, 25), ' ');
		std::cout << "  " << flags << o.name << ". Default value: " << o.get() << "\n";
	}
#if MLENS_CLI_STAR_FIELD
	std::cout << "Star-field options are ignored when a star file is given; -s is ignored\n"
				 "when -ks is given.\n";
#endif
#if !defined(MLENS_CLI_CCF) && !defined(MLENS_CLI_NCC)
	std::cout << "--precision (0 FP64, 1 mixed), --expansion_order (0 automatic) and\n"
				 "--m2l_mode (0-3) are this engine's additions.\n";
#endif
}

bool present(int argc, char** argv, const char* a, const char* b)
{
	for (int i = 1; i < argc; i++)
		if ((a[0] && std::string(argv[i]) == a) || std::string(argv[i]) == b) return true;
	return false;
}

}  // namespace

int main(int argc, char* argv[])
{
	if (present(argc, argv, "-h", "--help"))
	{
		usage(argv[0]);
		return -1;
	}
	if ((argc - 1) % 2 != 0)
	{
		std::cerr << "Error. Invalid input syntax.\n";
		usage(argv[0]);
		return -1;
	}
	const std::vector<Option> opts = options();
	auto find = [&](const std::string& flag) -> const Option* {
		for (const Option& o : opts)
			if ((o.shrt[0] && flag == o.shrt) || flag == o.lng) return &o;
		return nullptr;
	};
	// validate every option name, and read the verbosity first
	for (int i = 1; i < argc; i += 2)
	{
		const std::string flag = argv[i];
		if (flag == "-v" || flag == "--verbose")
		{
			try
			{
				verbose = std::stoi(argv[i + 1]);
			}
			catch (...)
			{
				std::cerr << "Error. Invalid verbose input.\n";
				return -1;
			}
			continue;
		}
		if (!find(flag))
		{
			std::cerr << "Error. Invalid input syntax. Unknown option " << flag << "\n";
			usage(argv[0]);
			return -1;
		}
	}
	const bool have_starfile = present(argc, argv, "-sf", "--starfile");
	const bool have_kappa_star = present(argc, argv, "-ks", "--kappa_star");
	[[maybe_unused]] const bool have_smooth = present(argc, argv, "-s", "--smooth_fraction");
	for (int i = 1; i < argc; i += 2)
	{
		const std::string flag = argv[i];
		if (flag == "-v" || flag == "--verbose") continue;
		const Option& o = *find(flag);
		if (o.star_field && have_starfile) continue;
		if (std::string(o.name) == "smooth_fraction" && have_kappa_star) continue;
		const std::string val = argv[i + 1];
		try
		{
			if (o.kind == REAL)
				o.set(std::stod(val), 0, val);
			else if (o.kind == INTEGER)
				o.set(0.0, std::stoi(val), val);
			else
				o.set(0.0, 0, val);
		}
		catch (...)
		{
			std::cerr << "Error. Invalid " << o.name << " input.\n";
			return -1;
		}
		if (verbose >= 2) std::cout << o.name << " set to: " << o.get() << "\n";
#if MLENS_CLI_STAR_FIELD
		if (std::string(o.name) == "smooth_fraction")
		{
			if (smooth_fraction < 0)
			{
				std::cerr << "Error. Invalid smooth_fraction input. smooth_fraction must be >= 0\n";
				return -1;
			}
			if (smooth_fraction >= 1)
			{
				std::cerr << "Error. Invalid smooth_fraction input. smooth_fraction must be < 1\n";
				return -1;
			}
		}
#endif
	}
#if MLENS_CLI_STAR_FIELD
	if (!have_starfile && !have_kappa_star && have_smooth)
	{
		maker.kappa_star = static_cast<dtype>((1 - smooth_fraction) * maker.kappa_tot);
		if (verbose >= 2) std::cout << "kappa_star set to: " << maker.kappa_star << "\n";
	}
#endif
	if (verbose >= 2) std::cout << "\n";
	if (!maker.run(verbose)) return -1;
	if (!maker.save(verbose)) return -1;
	return 0;
}
