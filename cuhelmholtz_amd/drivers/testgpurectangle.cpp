// testgpurectangle -- command-line driver with the arguments and output of the
// reference's testgpurectangle 
// (cudahelmholtz/testgpurectangle.cu:16-27): cells per axis and the three boundary codes.
#include <cstdlib>

#include "gpu_driver.h"

int main(int argc, char **argv)
{
    if (argc < 7) {
        printf("usage: ./testgpurectangle Nx Ny Nz xbc ybc zbc\n");
        return -1;
    }
    return drv::run_gpu("testgpurectangle", atoi(argv[1]), atoi(argv[2]), atoi(argv[3]), atoi(argv[4]), atoi(argv[5]),
                        atoi(argv[6]));
}
