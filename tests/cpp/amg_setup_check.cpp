// UA-AMG setup on the host alone (flecsolve_amd/csrc/fsx_amg_setup.hpp, the
// code fsx_amg_create runs): reads a CSR matrix, coarsens it as
// bound_solver::setup does and writes each level's fine-row -> aggregate
// map and coarse matrix, for comparison with the oracle's hierarchy.
//
//   amg_setup_check <dir> <beta> <passes> <max_levels> <min_coarse>
// reads <dir>/rp.bin, ci.bin (int64) and v.bin (float64); writes
// <dir>/aggt<l>.bin for each coarsened level l, <dir>/rp<l>.bin, ci<l>.bin,
// v<l>.bin for each coarse level l >= 1, and <dir>/levels.txt (the count).
#include "fsx_amg_setup.hpp"

#include <cstdio>
#include <string>

template<class T>
static std::vector<T> read_all(const std::string & path) {
	std::vector<T> out;
	FILE * f = std::fopen(path.c_str(), "rb");
	if (!f) return out;
	std::fseek(f, 0, SEEK_END);
	const long bytes = std::ftell(f);
	std::fseek(f, 0, SEEK_SET);
	out.resize((size_t)bytes / sizeof(T));
	if (!out.empty() && std::fread(out.data(), sizeof(T), out.size(), f) != out.size()) out.clear();
	std::fclose(f);
	return out;
}

template<class T>
static bool write_all(const std::string & path, const std::vector<T> & v) {
	FILE * f = std::fopen(path.c_str(), "wb");
	if (!f) return false;
	const bool ok = v.empty() || std::fwrite(v.data(), sizeof(T), v.size(), f) == v.size();
	return std::fclose(f) == 0 && ok;
}

int main(int argc, char ** argv) {
	if (argc != 6) {
		std::fprintf(stderr, "usage: amg_setup_check dir beta passes max_levels min_coarse\n");
		return 2;
	}
	const std::string dir = argv[1];
	const float beta = std::strtof(argv[2], nullptr);
	const int passes = std::atoi(argv[3]), max_levels = std::atoi(argv[4]);
	const int64_t min_coarse = std::atoll(argv[5]);
	fsx::hcsr H;
	H.rp = read_all<int64_t>(dir + "/rp.bin");
	H.ci = read_all<int64_t>(dir + "/ci.bin");
	H.v = read_all<double>(dir + "/v.bin");
	if (H.rp.empty() || H.ci.size() != H.v.size() || H.rp.back() != (int64_t)H.ci.size()) {
		std::fprintf(stderr, "bad input matrix\n");
		return 1;
	}
	H.n = (int64_t)H.rp.size() - 1;
	int levels = 1;
	bool ok = true;
	fsx::coarsen_levels(std::move(H), beta, passes, max_levels, min_coarse,
	                    [&](int i, const fsx::hcsr &, const fsx::aggregates &,
	                        const std::vector<int64_t> & aggt, const fsx::hcsr & C) {
		const std::string l = std::to_string(i), c = std::to_string(i + 1);
		ok = ok && write_all(dir + "/aggt" + l + ".bin", aggt) &&
		     write_all(dir + "/rp" + c + ".bin", C.rp) && write_all(dir + "/ci" + c + ".bin", C.ci) &&
		     write_all(dir + "/v" + c + ".bin", C.v);
		levels = i + 2;
	});
	FILE * f = std::fopen((dir + "/levels.txt").c_str(), "w");
	if (!f || !ok) return 1;
	std::fprintf(f, "%d\n", levels);
	std::fclose(f);
	return 0;
}
