// fmt_check.cpp -- TEST INFRASTRUCTURE ONLY.  The number formatting the
// reference's writers rely on, done the way they do it: std::ostringstream
// with setprecision(p) in the default float field (save_frame,
// rietkerk_bjork_FKE.cpp:2855-2861; the sweep file, :4866-4872).  The tests
// compare libtm's CSV text against it.
#include <cstring>
#include <iomanip>
#include <sstream>
#include <string>

extern "C" {

// ostream text of x at precision prec (prec <= 0: the stream default); returns
// the length, or -1 if cap is too small.
int fmt_ostream(double x, int prec, char* out, int cap)
{
    std::ostringstream oss;
    if (prec > 0) oss << std::setprecision(prec);
    oss << x;
    const std::string s = oss.str();
    if ((int)s.size() + 1 > cap) return -1;
    std::memcpy(out, s.c_str(), s.size() + 1);
    return (int)s.size();
}

}  // extern "C"
