// autogrid4 drop-in: the command line of setflags.cpp:43-193 on top of the
// C-ABI.  Success prints "<prog>: Successful Completion." to the log and
// exits 0; any failure logs "ERROR: ..." and "Unsuccessful Completion." and
// exits 1 (main.cpp:133-177, 3565).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <unistd.h>

#include "autogrid_amd.h"

static const char *kHelp =
    "-p parameter_filename\n"
    "\t\t-l log_filename\n"
    "\t\t-d (increment debug level)\n"
    "\t\t-h (display this message)\n"
    "\t\t--version (print version information, copyright, and license)\n"
    "\t\t--bhtree (the semantics of a -DUSE_BHTREE build)\n"
    "\t\t--binary-maps (also write raw FP32 maps and a binary .fld)\n";

int main(int argc, char **argv) {
    const char *prog = argv[0];
    const char *gpf = nullptr, *glg = nullptr;
    uint32_t flags = 0;
    for (int i = 1; i < argc; i++) {
        const char *a = argv[i];
        if (a[0] != '-') break;
        if (a[1] == '-') a++;
        switch (a[1]) {
            case 'd': break;
            case 'b':
                if (strcmp(a, "-bhtree") == 0) flags |= AG_JOB_BHTREE;
                else if (strcmp(a, "-binary-maps") == 0) flags |= AG_JOB_BINARY_MAPS;
                else { fprintf(stderr, "%s: unknown switch %s\n", prog, argv[i]); return 1; }
                break;
            case 'u':
            case 'h': printf("usage: AutoGrid %s\n", kHelp); return 0;
            case 'v':
                printf("AutoGrid %-8s\n", "4.2.7.x");
                printf("MI355X grid-map engine, C-ABI version %d, device kernels for gfx950\n", ag_abi_version());
                return 0;
            case 'l':
                if (i + 1 >= argc) { fprintf(stderr, "\n%s: Sorry, -l requires a filename.\n\t%s\n", prog, kHelp); return 1; }
                glg = argv[++i];
                break;
            case 'p':
                if (i + 1 >= argc) { fprintf(stderr, "\n%s: Sorry, -p requires a filename.\n\t%s\n", prog, kHelp); return 1; }
                gpf = argv[++i];
                break;
            default: fprintf(stderr, "%s: unknown switch -%c\n", prog, a[1]); return 1;
        }
    }
    if (!gpf && isatty(fileno(stdin))) {
        printf("usage: AutoGrid %s\n", kHelp);
        return 1;
    }
    // The library opens device 0 only after the GPF has been read and
    // checked, so input errors are reported the same way with or without a GPU.
    ag_status st = ag_run_gpf_flags(nullptr, gpf ? gpf : "", glg, prog, flags);
    return st == AG_OK ? 0 : 1;
}
