// Command-line handling shared by ASDMatch.out and prepare.out.  Options follow the reference's
// parser (utility/command_parser.cpp:12-32, utility/command.cpp:7-30): a flag is found by a linear
// scan of argv, its value is the next token ("" / 0 when absent); -a and -i are accepted and ignored.
#pragma once

#include <cstdlib>
#include <string>
#include <vector>

struct Args {
    std::vector<std::string> tok;
    Args(int argc, char **argv) {
        for (int i = 1; i < argc; ++i) tok.emplace_back(argv[i]);
    }
    std::string str(const char *flag) const {
        for (size_t i = 0; i + 1 < tok.size(); ++i)
            if (tok[i] == flag) return tok[i + 1];
        return "";
    }
    int num(const char *flag) const {
        std::string s = str(flag);
        return s.empty() ? 0 : std::atoi(s.c_str());
    }
};
