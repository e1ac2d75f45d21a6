// testgpu -- command-line driver with the arguments and output of the reference's
// testgpu (cudahelmholtz/testgpu.cu:15-30): N cells per axis and the three boundary codes.
#include <cstdlib>

#include "gpu_driver.h"

int main(int argc, char **argv)
{
    if (argc < 5) {
        printf("usage: ./testgpu N xbc ybc zbc\n");
        return -1;
    }
    const int N = atoi(argv[1]);
    return drv::run_gpu("testgpu", N, N, N, atoi(argv[2]), atoi(argv[3]), atoi(argv[4]));
}
