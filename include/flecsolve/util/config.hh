// flecsolve/util/config.hh -- configuration files for solver settings.
//
// Drop-in for the reference's read_config / with_label
// (flecsolve/util/config.hh:24-91), which parse boost::program_options
// config files; Boost is not a dependency here, so this header reads the
// same files itself with the semantics the reference relies on:
//   * `[section]` lines prefix the following keys with "section."; a line
//     `key = value` (surrounding blanks trimmed) sets option
//     "section.key"; `#` starts a comment; blank lines are skipped;
//   * options are typed: integers and floats must parse completely, bools
//     accept true/false, yes/no, on/off, 1/0 (any case), enums read with
//     their operator>>, std::vector<double> takes blank-separated numbers
//     (a multitoken option); a key given twice is an error;
//   * required options must be present, others take their default or, when
//     declared without one, keep the value they had;
//   * options may depend on earlier values (a factory adds the chosen
//     solver's options once its `type` is read), so the file is applied
//     repeatedly until the set of unrecognised keys stops changing (at most
//     50 passes, config.hh:57-81); unrecognised keys left at the end are an
//     error, as the reference's final strict parse makes them.
// Errors throw std::invalid_argument naming the file, key and problem.
#ifndef FLECSOLVE_AMD_UTIL_CONFIG_HH
#define FLECSOLVE_AMD_UTIL_CONFIG_HH

#include <algorithm>
#include <cctype>
#include <fstream>
#include <functional>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

namespace flecsolve {

struct with_label {
	with_label(const char * pre) : prefix(pre) {}
	void set_prefix(const char * pre) { prefix = pre; }

protected:
	std::string prefix;
	std::string label(const char * suf) { return {prefix + "." + suf}; }
};

namespace config {

template<class T>
T parse_value(const std::string & key, const std::string & text) {
	if constexpr (std::is_same_v<T, std::string>)
		return text;
	else if constexpr (std::is_same_v<T, std::vector<double>>) {
		// a multi-token value: blank-separated numbers
		std::istringstream is(text);
		std::vector<double> v;
		std::string tok;
		while (is >> tok) v.push_back(parse_value<double>(key, tok));
		if (v.empty())
			throw std::invalid_argument("option '" + key + "': empty list");
		return v;
	}
	else if constexpr (std::is_same_v<T, bool>) {
		std::string t(text);
		std::transform(t.begin(), t.end(), t.begin(),
		               [](unsigned char c) { return (char)std::tolower(c); });
		if (t == "true" || t == "yes" || t == "on" || t == "1") return true;
		if (t == "false" || t == "no" || t == "off" || t == "0") return false;
		throw std::invalid_argument("option '" + key + "': invalid bool '" + text + "'");
	}
	else {
		std::istringstream is(text);
		T v{};
		is >> v;
		if (is.fail() || !(is >> std::ws).eof())
			throw std::invalid_argument("option '" + key + "': invalid value '" + text + "'");
		return v;
	}
}

// One option: its full name, a setter from text, and what happens when the
// file does not give it (default or required).
struct option {
	std::string name;
	std::function<void(const std::string &)> set;
	std::function<void()> set_default; // empty: required
};

}

// The set of options one settings object accepts (the role of
// boost::program_options::options_description in the reference).
class optdesc {
public:
	template<class T>
	optdesc & required(const std::string & name, T * dst,
	                   std::function<void(const T &)> notify = {}) {
		opts_.push_back({name,
		                 [=](const std::string & v) {
			                 *dst = config::parse_value<T>(name, v);
			                 if (notify) notify(*dst);
		                 },
		                 {}});
		return *this;
	}
	template<class T, class D>
	optdesc & with_default(const std::string & name, T * dst, D def) {
		opts_.push_back({name,
		                 [=](const std::string & v) { *dst = config::parse_value<T>(name, v); },
		                 [=] { *dst = (T)def; }});
		return *this;
	}
	// no default: left as it is when the file does not give it (an option
	// declared without default_value in program_options)
	template<class T>
	optdesc & optional(const std::string & name, T * dst) {
		opts_.push_back({name,
		                 [=](const std::string & v) { *dst = config::parse_value<T>(name, v); },
		                 [] {}});
		return *this;
	}
	optdesc & add(const optdesc & o) {
		opts_.insert(opts_.end(), o.opts_.begin(), o.opts_.end());
		return *this;
	}
	const std::vector<config::option> & options() const { return opts_; }

private:
	std::vector<config::option> opts_;
};

namespace config {

// key -> value of a config file, in file order
inline std::vector<std::pair<std::string, std::string>> read_file(const char * fname) {
	std::ifstream in(fname);
	if (!in) throw std::invalid_argument(std::string("cannot open config file ") + fname);
	auto trim = [](std::string s) {
		const auto b = s.find_first_not_of(" \t\r\n");
		if (b == std::string::npos) return std::string();
		const auto e = s.find_last_not_of(" \t\r\n");
		return s.substr(b, e - b + 1);
	};
	std::vector<std::pair<std::string, std::string>> kv;
	std::string line, section;
	int lineno = 0;
	while (std::getline(in, line)) {
		++lineno;
		const auto hash = line.find('#');
		if (hash != std::string::npos) line.erase(hash);
		line = trim(line);
		if (line.empty()) continue;
		if (line.front() == '[') {
			if (line.back() != ']')
				throw std::invalid_argument(std::string(fname) + ":" + std::to_string(lineno) +
				                            ": bad section line");
			section = trim(line.substr(1, line.size() - 2));
			continue;
		}
		const auto eq = line.find('=');
		if (eq == std::string::npos)
			throw std::invalid_argument(std::string(fname) + ":" + std::to_string(lineno) +
			                            ": expected key = value");
		const std::string key = trim(line.substr(0, eq));
		const std::string name = section.empty() ? key : section + "." + key;
		for (const auto & p : kv)
			if (p.first == name)
				throw std::invalid_argument(std::string(fname) + ": option '" + name +
				                            "' given more than once");
		kv.emplace_back(name, trim(line.substr(eq + 1)));
	}
	return kv;
}

// One pass: apply every option of `desc`; returns the unrecognised keys.
inline std::vector<std::string>
apply(const char * fname, const std::vector<std::pair<std::string, std::string>> & kv,
      const optdesc & desc) {
	std::map<std::string, const option *> byname;
	for (const auto & o : desc.options()) byname[o.name] = &o;
	for (const auto & o : desc.options()) {
		auto it = std::find_if(kv.begin(), kv.end(), [&](const auto & p) { return p.first == o.name; });
		if (it != kv.end())
			o.set(it->second);
		else if (o.set_default)
			o.set_default();
		else
			throw std::invalid_argument(std::string(fname) + ": required option '" + o.name +
			                            "' is missing");
	}
	std::vector<std::string> unknown;
	for (const auto & p : kv)
		if (!byname.count(p.first)) unknown.push_back(p.first);
	return unknown;
}

}

// read_config (config.hh:47-91): one settings object per options object;
// a tuple of them for several, the object itself for one.
template<class... Options>
auto read_config(const char * fname, Options &&... ops) {
	std::tuple<typename std::decay_t<Options>::settings_type...> settings;
	const auto kv = config::read_file(fname);
	std::vector<std::string> prev{"-1"};
	std::vector<std::string> unknown;
	for (int depth = 0; depth <= 50; ++depth) {
		optdesc desc;
		std::apply([&](auto &... s) { (desc.add(ops(s)), ...); }, settings);
		unknown = config::apply(fname, kv, desc);
		if (unknown == prev) break;
		prev = unknown;
	}
	if (!unknown.empty())
		throw std::invalid_argument(std::string(fname) + ": unrecognised option '" +
		                            unknown.front() + "'");
	if constexpr (sizeof...(Options) == 1)
		return std::get<0>(settings);
	else
		return settings;
}

}

#endif
