// cli.cpp -- the umcx command-line program (main of src/umcx.cpp:868-875) on libumcx.so.
#include <iostream>

#include "../../include/umcx.h"

int main(int argn, char* argv[]) {
    int rc = MCX_run_simulation(argv, argn);
    if (rc != 0) {
        std::cout << "umcx encounters an error:" << std::endl << umcx_last_error() << std::endl;
        return 1;
    }
    return 0;
}
