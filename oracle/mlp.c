/* radiance-cache oracle */
