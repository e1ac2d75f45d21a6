// testcpu -- command-line driver with the arguments and output of the reference's
// testcpu (cudahelmholtz/testcpu.cpp:15-29): N cells per axis and the three boundary codes.
#include <cstdlib>

#include "cpu_driver.h"

int main(int argc, char **argv)
{
    if (argc < 5) {
        printf("usage: ./testcpu N xbc ybc zbc\n");
        return -1;
    }
    const int N = atoi(argv[1]);
    return drv::run_host("testcpu", N, N, N, atoi(argv[2]), atoi(argv[3]), atoi(argv[4]));
}
