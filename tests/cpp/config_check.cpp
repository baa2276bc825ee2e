// Config-file reading (read_config, solver options, krylov_factory options)
// without a device: parsed settings and the error cases.  Exit status 0 =
// pass.  argv[1] = the tests/golden directory.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <stdexcept>
#include <string>

#include "flecsolve/solvers/factory.hh"

using namespace flecsolve;

static int fails = 0;
#define CHECK(c)                                                       \
	do {                                                               \
		if (!(c)) {                                                    \
			std::printf("FAIL line %d: %s\n", __LINE__, #c);           \
			++fails;                                                   \
		}                                                              \
	} while (0)

static std::string tmpdir() {
	const char * t = std::getenv("TMPDIR");
	return t && *t ? t : "/tmp";
}

static bool throws(const std::string & text) {
	const std::string f = tmpdir() + "/fsx_config_check.cfg";
	{
		std::ofstream o(f);
		o << text;
	}
	try {
		(void)read_config(f.c_str(), krylov_factory::options("linear-solver"));
	}
	catch (const std::invalid_argument & e) {
		std::printf("  expected error: %s\n", e.what());
		std::remove(f.c_str());
		return true;
	}
	std::remove(f.c_str());
	return false;
}

int main(int argc, char ** argv) {
	const std::string dir = argc > 1 ? argv[1] : "tests/golden";
	{
		auto s = read_config((dir + "/krylov_cg.cfg").c_str(),
		                     krylov_factory::options("linear-solver"));
		CHECK(s.target_id && *s.target_id == krylov_target::cg);
		const auto & c = std::get<cg::settings>(s.target_settings);
		CHECK(c.maxiter == 10000 && !c.use_zero_guess && c.rtol == 1e-6f && c.atol == 0.f);
	}
	{
		auto s = read_config((dir + "/krylov_gmres.cfg").c_str(),
		                     krylov_factory::options("linear-solver"));
		CHECK(s.target_id && *s.target_id == krylov_target::gmres);
		const auto & g = std::get<gmres::settings>(s.target_settings);
		CHECK(g.maxiter == 3000 && g.use_zero_guess && g.rtol == 1e-10f);
		CHECK(g.max_krylov_dim == 30 && g.restart && g.pre_side == gmres::precond_side::left);
	}
	{
		auto s = read_config((dir + "/krylov_bicgstab.cfg").c_str(),
		                     krylov_factory::options("linear-solver"));
		CHECK(s.target_id && *s.target_id == krylov_target::bicgstab);
		const auto & b = std::get<bicgstab::settings>(s.target_settings);
		CHECK(b.maxiter == 400 && b.use_zero_guess && b.rtol == 1e-10f);
	}
	{
		// plain solver options under their own label; defaults applied
		const std::string f = tmpdir() + "/fsx_config_check_plain.cfg";
		{
			std::ofstream o(f);
			o << "[cg-solver]\nmaxiter = 7\nuse-zero-guess = off\n";
		}
		auto s = read_config(f.c_str(), cg::options("cg-solver"));
		std::remove(f.c_str());
		CHECK(s.maxiter == 7 && !s.use_zero_guess && s.rtol == 0.f && s.atol == 0.f);
	}
	// errors: missing required key, unknown key, bad values, unknown type,
	// repeated key, missing type
	CHECK(throws("[linear-solver]\ntype = cg\n[linear-solver.options]\nrtol = 1\n"
	             "use-zero-guess = true\n"));
	CHECK(throws("[linear-solver]\ntype = cg\n[linear-solver.options]\nmaxiter = 5\n"
	             "use-zero-guess = true\nmax-krylov-dim = 3\n"));
	CHECK(throws("[linear-solver]\ntype = cg\n[linear-solver.options]\nmaxiter = 5x\n"
	             "use-zero-guess = true\n"));
	CHECK(throws("[linear-solver]\ntype = cg\n[linear-solver.options]\nmaxiter = 5\n"
	             "use-zero-guess = maybe\n"));
	CHECK(throws("[linear-solver]\ntype = nka\n"));
	CHECK(throws("[linear-solver]\ntype = cg\ntype = gmres\n"));
	CHECK(throws("[linear-solver.options]\nmaxiter = 5\n"));
	CHECK(throws("[linear-solver]\ntype = gmres\n[linear-solver.options]\nmaxiter = 5\n"
	             "use-zero-guess = true\npre-side = middle\n"));
	std::printf(fails ? "FAIL\n" : "PASS\n");
	return fails ? 1 : 0;
}
