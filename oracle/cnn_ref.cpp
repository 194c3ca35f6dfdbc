/* placeholder, filled with the CNN FP64 oracle */
