// hq_rtc -- compile one specialised kernel source to a gfx950 code object.
//
//   hq_rtc IN.hip OUT.co LOG
//
// hiprtc compiles one program at a time per process (its compiler state is
// process-global), so the library runs concurrent compiles as child
// processes of this helper.  It touches no GPU.  Exit status: 0 compiled
// (OUT.co written), 1 the source did not compile (LOG holds the compiler
// log), 2 a file could not be read or written, 3 hiprtc itself failed.
#include "jit_opts.h"

#include <hip/hiprtc.h>

#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

static bool write_all(const char* path, const char* data, size_t n)
{
    std::ofstream f(path, std::ios::binary);
    if (!f) return false;
    f.write(data, (std::streamsize)n);
    f.close();
    return !f.fail();
}

int main(int argc, char** argv)
{
    if (argc != 4) {
        fprintf(stderr, "usage: hq_rtc IN.hip OUT.co LOG\n");
        return 2;
    }
    std::ifstream in(argv[1], std::ios::binary);
    if (!in) return 2;
    std::stringstream ss;
    ss << in.rdbuf();
    const std::string src = ss.str();
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, src.c_str(), "hq_group.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS) return 3;
    const hiprtcResult rc = hiprtcCompileProgram(prog, hq::kNumJitOpts, hq::kJitOpts);
    if (rc != HIPRTC_SUCCESS) {
        size_t n = 0;
        hiprtcGetProgramLogSize(prog, &n);
        std::string log(n, '\0');
        if (n) hiprtcGetProgramLog(prog, &log[0]);
        log = std::string(hiprtcGetErrorString(rc)) + "\n" + log;
        hiprtcDestroyProgram(&prog);
        return write_all(argv[3], log.data(), log.size()) ? 1 : 2;
    }
    size_t sz = 0;
    if (hiprtcGetCodeSize(prog, &sz) != HIPRTC_SUCCESS || sz == 0) {
        hiprtcDestroyProgram(&prog);
        return 3;
    }
    std::vector<char> code(sz);
    if (hiprtcGetCode(prog, code.data()) != HIPRTC_SUCCESS) {
        hiprtcDestroyProgram(&prog);
        return 3;
    }
    hiprtcDestroyProgram(&prog);
    return write_all(argv[2], code.data(), code.size()) ? 0 : 2;
}
