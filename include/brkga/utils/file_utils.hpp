// brkga/utils/file_utils.hpp - file helpers of the C++ facade.
//
// The helper names the reference's configs call (utils/file_utils.hpp):
// existence checks, path splitting, line reading and CSV field splitting.
#pragma once

#include <fstream>
#include <stdexcept>
#include <string>
#include <sys/stat.h>
#include <vector>

class FileUtils {
public:
    static bool file_exists(const std::string& path)
    {
        struct stat st;
        return ::stat(path.c_str(), &st) == 0 && S_ISREG(st.st_mode);
    }

    static bool directory_exists(const std::string& path)
    {
        struct stat st;
        return ::stat(path.c_str(), &st) == 0 && S_ISDIR(st.st_mode);
    }

    // Creates `path` and its missing parents.
    static void ensure_directory(const std::string& path)
    {
        if (path.empty() || directory_exists(path)) return;
        const std::size_t cut = path.find_last_of('/');
        if (cut != std::string::npos && cut > 0) ensure_directory(path.substr(0, cut));
        if (::mkdir(path.c_str(), 0755) != 0 && !directory_exists(path))
            throw std::runtime_error("Failed to create directory: " + path);
    }

    // Text after the last '.', or "" (utils/file_utils.hpp get_extension).
    static std::string get_extension(const std::string& path)
    {
        const std::size_t dot = path.find_last_of('.');
        return dot == std::string::npos ? std::string() : path.substr(dot + 1);
    }

    // File name without directory and without its last extension.
    static std::string get_basename(const std::string& path)
    {
        const std::size_t slash = path.find_last_of("/\\");
        std::string name = slash == std::string::npos ? path : path.substr(slash + 1);
        const std::size_t dot = name.find_last_of('.');
        return dot == std::string::npos ? name : name.substr(0, dot);
    }

    // Directory part, "." when there is none.
    static std::string get_directory(const std::string& path)
    {
        const std::size_t slash = path.find_last_of("/\\");
        return slash == std::string::npos ? std::string(".") : path.substr(0, slash);
    }

    // All lines, without their line terminators (\n or \r\n).
    static std::vector<std::string> read_lines(const std::string& path)
    {
        std::ifstream in(path);
        if (!in) throw std::runtime_error("Cannot open file: " + path);
        std::vector<std::string> lines;
        std::string line;
        while (std::getline(in, line)) {
            if (!line.empty() && line.back() == '\r') line.pop_back();
            lines.push_back(line);
        }
        return lines;
    }

    // Fields split on `delimiter`, each trimmed of blanks; a delimiter at
    // the very end does not open another field (getline splitting).
    static std::vector<std::string> parse_csv_line(const std::string& line, char delimiter = ',')
    {
        std::vector<std::string> out;
        std::size_t start = 0;
        while (start < line.size()) {
            std::size_t end = line.find(delimiter, start);
            if (end == std::string::npos) end = line.size();
            const std::string f = line.substr(start, end - start);
            const std::size_t a = f.find_first_not_of(" \t\r\n");
            const std::size_t b = f.find_last_not_of(" \t\r\n");
            out.push_back(a == std::string::npos ? std::string() : f.substr(a, b - a + 1));
            start = end + 1;
        }
        return out;
    }
};
