// Error types of the GooFit API (include/goofit/Error.h of the reference, where
// GeneralError derives from CLI::Error; here from std::runtime_error, with the
// same "{}" placeholder formatting of the message).
#pragma once

#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>

namespace GooFit {

namespace detail {
inline void format_into(std::ostringstream &os, const std::string &fmt, size_t pos) { os << fmt.substr(pos); }
template <typename T, typename... Args>
void format_into(std::ostringstream &os, const std::string &fmt, size_t pos, T &&v, Args &&...rest) {
    size_t at = fmt.find("{}", pos);
    if(at == std::string::npos) {
        os << fmt.substr(pos);
        return;
    }
    os << fmt.substr(pos, at - pos) << v;
    format_into(os, fmt, at + 2, std::forward<Args>(rest)...);
}
template <typename... Args>
std::string format(const std::string &fmt, Args &&...args) {
    std::ostringstream os;
    format_into(os, fmt, 0, std::forward<Args>(args)...);
    return os.str();
}
}  // namespace detail

/// Thrown when a general error is encountered
struct GeneralError : public std::runtime_error {
    template <typename... Args>
    explicit GeneralError(const std::string &msg, Args &&...args)
        : std::runtime_error(detail::format(msg, std::forward<Args>(args)...)) {}
    int get_exit_code() const { return 2; }
    std::string get_name() const { return "GeneralError"; }
};

struct OutOfRange : public GeneralError {
    OutOfRange(std::string name, double value, double min, double max)
        : GeneralError("{} value {} is not in range {} to {}", name, value, min, max) {}
};

/// Print a message with the parameter values and throw GeneralError
/// (GooFit::abort, src/goofit/Abort.cpp:10-42).
class PdfBase;
[[noreturn]] void abort(std::string file, int line, std::string reason, const PdfBase *pdf = nullptr);

}  // namespace GooFit
