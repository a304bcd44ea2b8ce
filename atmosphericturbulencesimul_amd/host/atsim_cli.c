/* atsim_cli.c -- command-line program: runs the commands given as arguments
 * (separated by ';') or, with none, one command per line of standard input,
 * as the reference's interpreter reads a script piped into it. */
#include <stdio.h>
#include <string.h>

#include "atsim_compat.h"

int main(int argc, char **argv) {
    int status = 0;
    atsim_compat_exit_on_error(1);
    if (argc > 1) {
        char line[4096] = "";
        for (int i = 1; i < argc; i++) {
            strncat(line, argv[i], sizeof line - strlen(line) - 2);
            strcat(line, " ");
        }
        char *save = NULL;
        for (char *cmd = strtok_r(line, ";", &save); cmd; cmd = strtok_r(NULL, ";", &save)) {
            int rc = atsim_cli_execute(cmd);
            if (rc > 0) break;
            if (rc < 0) status = 1;
        }
        return status;
    }
    char line[4096];
    while (fgets(line, sizeof line, stdin)) {
        int rc = atsim_cli_execute(line);
        if (rc > 0) break;
        if (rc < 0) status = 1;
    }
    return status;
}
