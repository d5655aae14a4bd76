// app_resume.h -- resuming an interrupted run of the executables from their own output file.
//
// The executables write their output in units (a batch of triples, an r bin of triangles) and flush each
// unit before computing the next.  With --resume they read the output file back, keep the lines of the
// units it already holds in full (a unit cut off by an interrupted run, or a last line without its newline,
// is dropped and computed again) and append from there.  No counterpart in the reference, whose
// executables write everything at the end.
#ifndef TP_APP_RESUME_H
#define TP_APP_RESUME_H
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

namespace tp_app {

// the lines of path that end with a newline (none if the file does not exist)
inline std::vector<std::string> complete_lines(const std::string &path)
{
    std::vector<std::string> out;
    std::ifstream in(path, std::ios::binary);
    if (!in) return out;
    std::stringstream ss;
    ss << in.rdbuf();
    const std::string s = ss.str();
    size_t a = 0;
    for (size_t b = s.find('\n'); b != std::string::npos; a = b + 1, b = s.find('\n', a)) out.push_back(s.substr(a, b - a));
    return out;
}

// rewrites path with its first n complete lines and returns a stream that appends after them
inline std::ofstream keep_and_append(const std::string &path, const std::vector<std::string> &lines, size_t n)
{
    {
        std::ofstream t(path, std::ios::binary | std::ios::trunc);
        for (size_t i = 0; i < n && i < lines.size(); ++i) t << lines[i] << '\n';
    }
    return std::ofstream(path, std::ios::binary | std::ios::app);
}

// the first k whitespace-separated numbers of a line
inline std::vector<double> leading_numbers(const std::string &line, size_t k)
{
    std::istringstream ss(line);
    std::vector<double> v;
    double x;
    while (v.size() < k && ss >> x) v.push_back(x);
    return v;
}

}  // namespace tp_app
#endif
